"""C ABI checks that need no GPU: the library loads, exports every entry
point include/hdc_amd.h declares, validates arguments before touching the
device, and its host-side packing / synthetic generator behave."""
import ctypes
import os
import re

import numpy as np
import pytest

import hdc_amd as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "hdc_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hvk_\w+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    L = H.lib()
    names = header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert sorted(H.EXPORTS) == names


def test_status_strings():
    L = H.lib()
    assert L.hvk_status_string(0) == b"ok"
    assert b"argument" in L.hvk_status_string(1)
    assert L.hvk_status_string(99) == b"unknown status"


def test_ctx_create_rejects_bad_args_without_device():
    L = H.lib()
    h = ctypes.c_void_p()
    assert L.hvk_ctx_create(32, 5, ctypes.byref(h)) == H.HVK_ERR_ARG
    assert L.hvk_ctx_create(10240, 0, ctypes.byref(h)) == H.HVK_ERR_ARG
    assert L.hvk_ctx_create(10240, 17, ctypes.byref(h)) == H.HVK_ERR_ARG
    assert L.hvk_ctx_create(40000, 5, ctypes.byref(h)) == H.HVK_ERR_ARG
    assert L.hvk_ctx_create(10240, 12, None) == H.HVK_ERR_ARG


def test_null_context_is_an_argument_error():
    L = H.lib()
    assert L.hvk_encode(None, None, None, None, 1, 10, None, None, None, 0, None, None, None,
                        None) == H.HVK_ERR_ARG
    assert L.hvk_search_plain(None, None, None, 1, 1, None, None, 1, 1, 0, 5, None, None, None,
                              None, None) == H.HVK_ERR_ARG
    assert L.hvk_search_mfma(None, None, None, None, 1, None, None, None, 1, 0, 5, None, None,
                             None, None, None) == H.HVK_ERR_ARG
    assert L.hvk_ctx_destroy(None) == H.HVK_ERR_ARG
    assert L.hvk_ctx_set_timing(None, 1) == H.HVK_ERR_ARG
    assert L.hvk_ctx_set_tuning(None, H.TUNE_SPLITS, 8) == H.HVK_ERR_ARG


def test_pack_layout():
    seqs = ["ACGTNacgtRY", "", "TTTTGGGGCCCCAAAAT", "A" * 40]
    buf, off = H.concat(seqs)
    packed, invalid = H.pack(buf, off)
    s = buf.decode()
    assert len(packed) == len(s) // 16 + 2
    code = {"A": 0, "C": 1, "G": 2, "T": 3}
    for p, ch in enumerate(s):
        c = (int(packed[p // 16]) >> (2 * (p % 16))) & 3
        bad = (int(invalid[p // 32]) >> (p % 32)) & 1
        assert bad == (0 if ch in code else 1)
        assert c == code.get(ch, 0)


def test_pack_rejects_decreasing_offsets():
    L = H.lib()
    buf = ctypes.create_string_buffer(b"ACGT")
    off = np.array([0, 3, 2], np.int64)
    a = np.zeros(4, np.uint32)
    assert L.hvk_pack(buf, off.ctypes.data, 2, a.ctypes.data, a.ctypes.data) == H.HVK_ERR_ARG


def test_generator_deterministic_and_in_range():
    spec = H.make_spec(seed=3, n_ancestors=5, desc_per_ancestor=6, anc_len_min=1300,
                       anc_len_max=1600, div_min=0.02, div_max=0.08, n_reads=200,
                       read_len_min=150, read_len_max=300, err_min=0.005, err_max=0.02)
    a = H.generate(spec)
    b = H.generate(spec)
    assert a.refs == b.refs and a.reads == b.reads and np.array_equal(a.truth, b.truth)
    c = H.generate(H.make_spec(seed=4, n_reads=10))
    assert c.refs != a.refs
    rl = np.diff(a.ref_offsets)
    assert len(rl) == 30 and rl.min() >= 1300 * 0.95 and rl.max() <= 1600 * 1.05
    ql = np.diff(a.read_offsets)
    assert ql.min() >= 150 and ql.max() <= 300
    assert set(a.refs.decode()) <= set("ACGT") and set(a.reads.decode()) <= set("ACGT")
    # every read is a window of its source reference with a few substitutions
    for i in range(0, 200, 17):
        r = a.reads[a.read_offsets[i]:a.read_offsets[i + 1]]
        src = a.refs[a.ref_offsets[a.truth[i]]:a.ref_offsets[a.truth[i] + 1]]
        best = min(sum(x != y for x, y in zip(r, src[s:s + len(r)]))
                   for s in range(len(src) - len(r) + 1))
        assert best <= max(3, int(0.06 * len(r)))


def test_generator_log_uniform_lengths():
    spec = H.make_spec(n_ancestors=10, desc_per_ancestor=10, anc_len_min=1300,
                       anc_len_max=1600, n_reads=3000, read_len_min=100, read_len_max=1600,
                       read_len_log_uniform=1)
    w = H.generate(spec)
    ql = np.diff(w.read_offsets)
    assert ql.min() >= 100 and ql.max() <= 1600
    # log-uniform: about as many reads in [100,400) as in [400,1600)
    lo = (ql < 400).mean()
    assert 0.4 < lo < 0.6


def test_generator_rejects_bad_spec():
    L = H.lib()
    bad = H.make_spec(read_len_min=300, read_len_max=100)
    a, b = ctypes.c_int64(), ctypes.c_int64()
    assert L.hvk_gen_size(ctypes.byref(bad), ctypes.byref(a), ctypes.byref(b)) == H.HVK_ERR_ARG


def test_cli_usage_errors_throw():
    """The drop-in tools reject a wrong argument count as the reference does
    (std::runtime_error with its usage text), before touching the device."""
    import subprocess
    b = os.path.join(ROOT, "bacterial-classification_amd", "bin")
    if not os.path.exists(os.path.join(b, "identify")):
        pytest.skip("tools not built")
    p = subprocess.run([os.path.join(b, "identify"), "a"], capture_output=True, text=True)
    assert p.returncode != 0 and "Usage: identify" in p.stderr
    p = subprocess.run([os.path.join(b, "propogate-children")], capture_output=True, text=True)
    assert p.returncode != 0 and "Usage: propogate-children" in p.stderr


def test_calib_rejects_bad_args_without_device():
    L = H.lib()
    sink = ctypes.c_void_p(16)  # never dereferenced: every call below fails validation first
    ops = ctypes.c_double()
    for wps, iters, qs, rs, snk in ((0, 1, 15, 39, sink), (3, 1, 15, 39, sink), (1, 0, 15, 39, sink),
                                    (1, 1, -1, 39, sink), (1, 1, 15, 128, sink), (2, 8, 15, 39, None),
                                    (1, (1 << 24) + 1, 15, 39, sink)):
        assert L.hvk_calib_mfma_i8(wps, iters, qs, rs, snk, ctypes.byref(ops), None) == H.HVK_ERR_ARG
