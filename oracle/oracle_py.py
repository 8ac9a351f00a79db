"""ctypes loader of the C oracle (liboracle.so) and of the reference build.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Never part of the product.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

_o = None
_r = None


def oracle():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"{ORACLE_SO} missing: make -C oracle")
        L = ctypes.CDLL(ORACLE_SO)
        vp, i32, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
        L.orc_mt19937_64.argtypes = [ctypes.c_uint64, vp, sz]
        L.orc_words.argtypes = [i32]
        L.orc_empty_hv.argtypes = [i32, vp]
        L.orc_base_hv.argtypes = [i32, ctypes.c_char, vp]
        L.orc_mult.argtypes = [i32, vp, vp, vp, sz]
        L.orc_kmer_hv.argtypes = [i32, i32, ctypes.c_char_p, vp]
        L.orc_superpose_batch.argtypes = [i32, vp, vp, sz]
        L.orc_replace_batch.argtypes = [i32, vp, vp, vp, sz]
        L.orc_compute.argtypes = [i32, i32, ctypes.c_char_p, sz, vp]
        L.orc_compute32.argtypes = [i32, i32, ctypes.c_char_p, sz, vp]
        L.orc_cosine.restype = ctypes.c_float
        L.orc_cosine.argtypes = [i32, vp, vp, i32]
        L.orc_dot16.restype = ctypes.c_int64
        L.orc_dot16.argtypes = [i32, vp, vp]
        L.orc_cos64_from.restype = ctypes.c_double
        L.orc_cos64_from.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]
        L.orc_topk16.argtypes = [i32, vp, i32, vp, i32, i32, vp, vp, i32]
        L.orc_precompute_similarity.restype = ctypes.c_float
        L.orc_precompute_similarity.argtypes = [i32, i32, i32, i32, ctypes.c_char_p, sz,
                                                ctypes.c_char_p, sz, ctypes.POINTER(sz)]
        L.orc_precompute_windows.argtypes = [i32, i32, i32, i32, ctypes.c_char_p, sz, vp, i32]
        L.orc_descend.argtypes = [i32, vp, vp, vp, vp, i32, vp, i32, i32, ctypes.c_float, i32,
                                  vp, vp, vp, vp, i32]
        L.orc_descend_scores.argtypes = [i32, vp, vp, vp, i32, vp]
        L.orc_tax_order.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i32, vp, vp, i32, vp,
                                    ctypes.POINTER(ctypes.c_int)]
        L.orc_classify_reads.restype = ctypes.c_double
        L.orc_classify_reads.argtypes = [i32, i32, ctypes.c_char_p, vp, i32, vp, vp, i32, i32,
                                         vp, vp, i32]
        _o = L
    return _o


def reference():
    """The reference's own CPU sources built by oracle/Makefile (or None)."""
    global _r
    if _r is None:
        if not os.path.exists(REF_SO):
            return None
        L = ctypes.CDLL(REF_SO)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.ref_config.argtypes = [ctypes.POINTER(ctypes.c_int)] * 4
        L.ref_base_hv.argtypes = [ctypes.c_char, vp]
        L.ref_empty_hv.argtypes = [vp]
        L.ref_mult.argtypes = [vp, vp, vp, sz]
        L.ref_kmer_hv.argtypes = [ctypes.c_char_p, vp]
        L.ref_compute.argtypes = [ctypes.c_char_p, sz, vp]
        L.ref_superpose_batch.argtypes = [vp, vp, sz]
        L.ref_replace_batch.argtypes = [vp, vp, vp, sz]
        L.ref_cosine.restype = ctypes.c_float
        L.ref_cosine.argtypes = [vp, vp]
        L.ref_precompute_similarity.restype = ctypes.c_float
        L.ref_precompute_similarity.argtypes = [ctypes.c_char_p, sz, ctypes.c_char_p, sz,
                                                ctypes.POINTER(sz)]
        L.ref_precompute_windows.argtypes = [ctypes.c_char_p, sz, vp, ctypes.c_int]
        L.ref_kmers_all.argtypes = [vp]
        _r = L
    return _r


def ref_config():
    r = reference()
    vals = [ctypes.c_int() for _ in range(4)]
    r.ref_config(*[ctypes.byref(v) for v in vals])
    return tuple(v.value for v in vals)


# ---- convenience wrappers --------------------------------------------------
def words(D):
    return (D + 63) // 64


def mt19937_64(seed, n):
    out = np.zeros(n, np.uint64)
    oracle().orc_mt19937_64(seed, out.ctypes.data, n)
    return out


def base_hv(D, ch):
    out = np.zeros(words(D), np.uint64)
    oracle().orc_base_hv(D, ch.encode(), out.ctypes.data)
    return out


def kmer_hv(D, k, kmer):
    out = np.zeros(words(D), np.uint64)
    ok = oracle().orc_kmer_hv(D, k, kmer.encode(), out.ctypes.data)
    return out if ok else None


def compute(D, k, seq):
    b = seq.encode() if isinstance(seq, str) else seq
    out = np.zeros(D, np.int16)
    n = oracle().orc_compute(D, k, b, len(b), out.ctypes.data)
    return out, n


def compute32(D, k, seq):
    b = seq.encode() if isinstance(seq, str) else seq
    out = np.zeros(D, np.int32)
    n = oracle().orc_compute32(D, k, b, len(b), out.ctypes.data)
    return out, n


def cosine(D, l, r, norm_mode=1):
    l = np.ascontiguousarray(l, np.int16)
    r = np.ascontiguousarray(r, np.int16)
    return oracle().orc_cosine(D, l.ctypes.data, r.ctypes.data, norm_mode)


def topk16(q, db, k, threads=8):
    q = np.ascontiguousarray(q, np.int16)
    db = np.ascontiguousarray(db, np.int16)
    nq, D = q.shape
    idx = np.zeros((nq, k), np.int32)
    cos = np.zeros((nq, k), np.float64)
    oracle().orc_topk16(D, q.ctypes.data, nq, db.ctypes.data, db.shape[0], k, idx.ctypes.data,
                        cos.ctypes.data, threads)
    return idx, cos


def precompute_similarity(D, k, window, step, query, ref):
    q, r = query.encode(), ref.encode()
    pos = ctypes.c_size_t()
    s = oracle().orc_precompute_similarity(D, k, window, step, q, len(q), r, len(r),
                                           ctypes.byref(pos))
    return s, pos.value


def precompute_windows(D, k, window, step, ref, max_windows=4096):
    r = ref.encode()
    out = np.zeros((max_windows, D), np.int16)
    n = oracle().orc_precompute_windows(D, k, window, step, r, len(r), out.ctypes.data, max_windows)
    return out[:n]


def descend(proto, child_off, child_idx, roots, q, beam, thr, levels, threads=8):
    """identify's beam over per-category prototypes (orc_descend).
    proto [n_cat][D] int32, q [n][D] int16; returns path, pscore, len, exit."""
    proto = np.ascontiguousarray(proto, np.int32)
    q = np.ascontiguousarray(q, np.int16)
    co = np.ascontiguousarray(child_off, np.int32)
    ci = np.ascontiguousarray(child_idx, np.int32)
    ro = np.ascontiguousarray(roots, np.int32)
    n, D = q.shape
    path = np.full((n, levels), -1, np.int32)
    ps = np.zeros((n, levels), np.float32)
    ln = np.zeros(n, np.int32)
    ex = np.zeros(n, np.int32)
    oracle().orc_descend(D, proto.ctypes.data, co.ctypes.data, ci.ctypes.data, ro.ctypes.data,
                         len(ro), q.ctypes.data, n, beam, thr, levels, path.ctypes.data,
                         ps.ctypes.data, ln.ctypes.data, ex.ctypes.data, threads)
    return path, ps, ln, ex


def tax_order(tsv_path, names):
    """Candidate lists (child_off, child_idx) and roots in the order of the
    reference's Hierarchy containers (orc_tax_order)."""
    n = len(names)
    cap = 16 * n + 1024
    off = np.zeros(n + 1, np.int32)
    idx = np.zeros(cap, np.int32)
    roots = np.zeros(max(n, 1), np.int32)
    nr = ctypes.c_int()
    m = oracle().orc_tax_order(tsv_path.encode(), "\n".join(names).encode(), n, off.ctypes.data,
                               idx.ctypes.data, cap, roots.ctypes.data, ctypes.byref(nr))
    if m < 0:
        raise ValueError("candidate lists exceed capacity")
    return off, idx[:m].copy(), roots[: nr.value].copy()


# ---- CPU baseline build -----------------------------------------------------
RELEASE_FLAGS = ["-O3", "-march=native", "-ffast-math"]  # the reference's Release (src/CMakeLists.txt:28)


def release_classifier(out_dir):
    """orc_classify_reads (the C restatement of the reference's bit-packed CPU
    path) compiled here with the reference's Release flags for this host's
    CPU, into out_dir.  Returns (ctypes library, flags) or (None, error text)
    when gcc is unavailable or fails."""
    import subprocess
    so = os.path.join(out_dir, "liboracle_release.so")
    obj = os.path.join(out_dir, "hdc_oracle_release.o")
    try:
        subprocess.run(["gcc", *RELEASE_FLAGS, "-fPIC", "-c", os.path.join(HERE, "hdc_oracle.c"),
                        "-o", obj], check=True, capture_output=True, text=True)
        subprocess.run(["gcc", "-shared", "-o", so, obj, "-lpthread", "-lm"], check=True,
                       capture_output=True, text=True)
    except (OSError, subprocess.CalledProcessError) as e:
        return None, str(getattr(e, "stderr", e))[-500:]
    L = ctypes.CDLL(so)
    vp, i32 = ctypes.c_void_p, ctypes.c_int
    L.orc_classify_reads.restype = ctypes.c_double
    L.orc_classify_reads.argtypes = [i32, i32, ctypes.c_char_p, vp, i32, vp, vp, i32, i32, vp, vp,
                                     i32]
    return L, " ".join(RELEASE_FLAGS)


def host_cpu():
    """(nproc, cores this process may use, CPU model name)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return os.cpu_count() or 1, avail, model

