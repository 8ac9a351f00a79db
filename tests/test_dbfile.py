"""Prototype database file (hvk_db_*), host side: round trip of every array
and the names, row-stride repacking, partial array sets, and rejection of
corrupted, truncated or foreign files.  The device upload is covered by
tests/test_gpu_dbfile.py."""
import os

import numpy as np
import pytest


def _arrays(rng, n, ld, dim):
    counts = np.zeros((n, ld), np.int16)
    counts[:, :dim] = rng.integers(-1500, 1500, (n, dim))
    lo = (counts & 0xFF).astype(np.uint8).view(np.int8)
    hi = ((counts.astype(np.int32) - lo) >> 8).astype(np.int8)
    c8 = np.clip(counts, -127, 127).astype(np.int8)
    sq = (counts.astype(np.int64) ** 2).sum(1)
    mx = np.abs(counts).max(1).astype(np.int32)
    res = np.maximum(np.abs(counts.astype(np.int32)) - 127, 0).sum(1).astype(np.int32)
    nk = rng.integers(0, 2000, n).astype(np.int32)
    return dict(counts=counts, lo=lo, hi=hi, clamp8=c8, sqnorm=sq, maxabs=mx, resid1=res, nkmers=nk)


def test_round_trip(tmp_path):
    import hdc_amd as H
    rng = np.random.default_rng(3)
    n, ld, dim = 37, 256, 250
    a = _arrays(rng, n, ld, dim)
    names = b"\n".join(b"ref%d" % i for i in range(n))
    path = str(tmp_path / "db.hvk")
    H.db_write_host(path, dim, 7, ld, a, names)
    f = H.DbFile(path)
    assert (f.dim, f.kmer, f.ld, f.n, f.flags) == (dim, 7, ld, n, 31)
    assert np.array_equal(f.array(H.DB_COUNTS, np.int16, (n, ld)), a["counts"])
    assert np.array_equal(f.array(H.DB_LO, np.int8, (n, ld)), a["lo"])
    assert np.array_equal(f.array(H.DB_HI, np.int8, (n, ld)), a["hi"])
    assert np.array_equal(f.array(H.DB_CLAMP8, np.int8, (n, ld)), a["clamp8"])
    for i, (key, dt) in enumerate((("sqnorm", np.int64), ("maxabs", np.int32),
                                   ("resid1", np.int32), ("nkmers", np.int32))):
        assert np.array_equal(f.array(32 + i, dt, (n,)), a[key])
    assert f.names() == names
    f.close()
    # sections start on 4096-byte boundaries after the 128-byte header
    raw = open(path, "rb").read()
    assert raw[:8] == b"HVKPDB01" and len(raw) % 4096 == len(names) % 4096


def test_partial_arrays_and_repack(tmp_path):
    import hdc_amd as H
    rng = np.random.default_rng(4)
    n, dim = 9, 130
    a = _arrays(rng, n, 384, dim)
    path = str(tmp_path / "db2.hvk")
    H.db_write_host(path, dim, 5, 256, dict(lo=a["lo"], sqnorm=a["sqnorm"], maxabs=a["maxabs"],
                                            resid1=a["resid1"], nkmers=a["nkmers"]))
    f = H.DbFile(path)
    assert f.flags == H.DB_LO | H.DB_SUMS and f.ld == 256
    assert f.array(H.DB_COUNTS, np.int16, (n, 256)) is None
    assert np.array_equal(f.array(H.DB_LO, np.int8, (n, 256)), a["lo"][:, :256])
    assert f.names() == b""
    f.close()


def test_empty_database(tmp_path):
    import hdc_amd as H
    path = str(tmp_path / "empty.hvk")
    H.db_write_host(path, 64, 3, 128, {})
    f = H.DbFile(path)
    assert (f.n, f.flags) == (0, 0)
    f.close()


@pytest.mark.parametrize("where", ["magic", "version", "dim", "checksum", "truncate"])
def test_rejects_bad_files(tmp_path, where):
    import hdc_amd as H
    rng = np.random.default_rng(5)
    a = _arrays(rng, 5, 128, 100)
    path = str(tmp_path / "bad.hvk")
    H.db_write_host(path, 100, 5, 128, a)
    raw = bytearray(open(path, "rb").read())
    if where == "magic":
        raw[0] ^= 1
    elif where == "version":
        raw[8] = 2
    elif where == "dim":
        raw[16] ^= 4  # the header checksum no longer matches
    elif where == "checksum":
        raw[120] ^= 0x80
    else:
        raw = raw[: len(raw) - 4096]
    open(path, "wb").write(bytes(raw))
    with pytest.raises(H.HvkError):
        H.DbFile(path)


def test_write_rejects_missing_arrays(tmp_path):
    import hdc_amd as H
    import ctypes
    o = H.EncodeOut()
    rc = H.lib().hvk_db_write(str(tmp_path / "x.hvk").encode(), 100, 5, 128, 4, H.DB_LO,
                              ctypes.byref(o), None, 0)
    assert rc == H.HVK_ERR_ARG
    assert not os.path.exists(tmp_path / "x.hvk")
