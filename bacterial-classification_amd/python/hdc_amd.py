"""ctypes binding of libhdc_amd.so (the gfx950 HDC read-classification path).

Thin plumbing for tests and the benchmark: device buffers are torch tensors,
kernels run through the C ABI declared in include/hdc_amd.h.  There is no
fallback: if the library is missing or has no GPU, calls raise.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libhdc_amd.so")

HVK_OK = 0
HVK_ERR_ARG = 1
HVK_ERR_HIP = 2
HVK_ERR_RANGE = 3
HVK_ERR_NOMEM = 4
HVK_ERR_NODEVICE = 5

# entry points declared in include/hdc_amd.h
EXPORTS = [
    "hvk_status_string", "hvk_last_error", "hvk_ctx_create", "hvk_ctx_destroy",
    "hvk_ctx_info", "hvk_base_hv", "hvk_pack_words", "hvk_pack", "hvk_encode",
    "hvk_search_plain", "hvk_search_mfma", "hvk_gen_size", "hvk_gen_fill",
    "hvk_ctx_set_timing", "hvk_ctx_last_timing", "hvk_encode_ex", "hvk_encode_windows",
    "hvk_search_mfma_bounded", "hvk_prepare_counts", "hvk_ctx_set_tuning",
    "hvk_descend", "hvk_tree_place", "hvk_tree_sum", "hvk_tree_limbs",
    "hvk_batch_create", "hvk_batch_destroy", "hvk_batch_config", "hvk_batch_get_config",
    "hvk_batch_classify", "hvk_batch_autotune", "hvk_batch_descend",
    "hvk_db_write", "hvk_db_open", "hvk_db_close", "hvk_db_info", "hvk_db_array", "hvk_db_upload",
    "hvk_search_mfma_bounded_noalloc", "hvk_calib_mfma_i8",
]

DB_COUNTS, DB_LO, DB_HI, DB_CLAMP8, DB_SUMS = 1, 2, 4, 8, 16

BATCH_FORM_BOUNDED = 0
BATCH_FORM_TWO_LIMB = 1

TUNE_SPLITS = 1
TUNE_SELECT_KIB = 2
TUNE_SHAPE = 3
TUNE_ENCODER = 4
TUNE_DESCENT = 5
TUNE_GROUP_SLICES = 6


class HvkError(RuntimeError):
    def __init__(self, status: int, what: str, lib=None):
        msg = f"{what}: status {status}"
        if lib is not None:
            msg += f" ({lib.hvk_status_string(status).decode()}; {lib.hvk_last_error().decode()})"
        super().__init__(msg)
        self.status = status


class EncodeOut(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("counts", "lo", "hi", "clamp8")] + [
        ("ld", ctypes.c_int64)] + [(n, ctypes.c_void_p) for n in ("sqnorm", "maxabs", "nkmers",
                                                                  "resid1")]


class Bundles(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("clamp8", "lo", "hi", "sqnorm", "maxabs",
                                               "resid1")] + [("n", ctypes.c_int)]


class Tree(ctypes.Structure):
    _fields_ = [("n_cat", ctypes.c_int), ("ld", ctypes.c_int64), ("d_lo", ctypes.c_void_p),
                ("d_hi", ctypes.c_void_p), ("d_sqnorm", ctypes.c_void_p),
                ("max_abs_limb", ctypes.c_int32), ("child_off", ctypes.c_void_p),
                ("child_idx", ctypes.c_void_p), ("roots", ctypes.c_void_p),
                ("n_roots", ctypes.c_int), ("max_abs_proto", ctypes.c_int64),
                ("d_row_maxabs", ctypes.c_void_p)]


class GenSpec(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("n_ancestors", ctypes.c_int32),
        ("desc_per_ancestor", ctypes.c_int32),
        ("anc_len_min", ctypes.c_int32),
        ("anc_len_max", ctypes.c_int32),
        ("div_min", ctypes.c_double),
        ("div_max", ctypes.c_double),
        ("indel_rate", ctypes.c_double),
        ("n_reads", ctypes.c_int32),
        ("read_len_min", ctypes.c_int32),
        ("read_len_max", ctypes.c_int32),
        ("read_len_log_uniform", ctypes.c_int32),
        ("err_min", ctypes.c_double),
        ("err_max", ctypes.c_double),
    ]


_lib = None


def lib():
    """Load libhdc_amd.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} missing: run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    L.hvk_status_string.restype = ctypes.c_char_p
    L.hvk_status_string.argtypes = [i32]
    L.hvk_last_error.restype = ctypes.c_char_p
    L.hvk_last_error.argtypes = []
    L.hvk_ctx_create.argtypes = [i32, i32, ctypes.POINTER(vp)]
    L.hvk_ctx_destroy.argtypes = [vp]
    L.hvk_ctx_info.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int)] * 4
    L.hvk_base_hv.argtypes = [vp, ctypes.c_char, vp]
    L.hvk_pack_words.restype = i64
    L.hvk_pack_words.argtypes = [i64]
    L.hvk_pack.argtypes = [vp, vp, i32, vp, vp]
    L.hvk_encode.argtypes = [vp, vp, vp, vp, i32, i32, vp, vp, vp, i64, vp, vp, vp, vp]
    L.hvk_search_plain.argtypes = [vp, vp, vp, i32, i32, vp, vp, i32, i32, i64, i32,
                                   vp, vp, vp, vp, vp]
    L.hvk_search_mfma.argtypes = [vp, vp, vp, vp, i32, vp, vp, vp, i32, i64, i32,
                                  vp, vp, vp, vp, vp]
    L.hvk_encode_ex.argtypes = [vp, vp, vp, vp, i32, i32, ctypes.POINTER(EncodeOut), vp]
    L.hvk_encode_windows.argtypes = [vp, vp, vp, vp, i32, i32, ctypes.POINTER(EncodeOut), vp]
    L.hvk_search_mfma_bounded.argtypes = [vp, ctypes.POINTER(Bundles), ctypes.POINTER(Bundles),
                                          i64, i32, vp, vp, vp, vp, vp]
    L.hvk_search_mfma_bounded_noalloc.argtypes = L.hvk_search_mfma_bounded.argtypes
    L.hvk_prepare_counts.argtypes = [vp, vp, i32, ctypes.POINTER(EncodeOut), vp]
    L.hvk_ctx_set_timing.argtypes = [vp, i32]
    L.hvk_ctx_set_tuning.argtypes = [vp, i32, i32]
    L.hvk_ctx_last_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    L.hvk_gen_size.argtypes = [ctypes.POINTER(GenSpec), ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.hvk_gen_fill.argtypes = [ctypes.POINTER(GenSpec), vp, vp, vp, vp, vp]
    L.hvk_descend.argtypes = [vp, ctypes.POINTER(Tree), vp, vp, i32, i64, i32, i32,
                              ctypes.c_float, i32, vp, vp, vp, vp, vp]
    L.hvk_tree_place.argtypes = [vp, vp, i64, i32, vp, vp, i64, i32, vp]
    L.hvk_tree_sum.argtypes = [vp, vp, i64, i32, vp, vp, vp, i32, vp, vp]
    L.hvk_tree_limbs.argtypes = [vp, vp, i32, i64, vp, vp, vp, vp, vp, vp]
    L.hvk_db_write.argtypes = [ctypes.c_char_p, i32, i32, i64, i32, ctypes.c_uint32,
                               ctypes.POINTER(EncodeOut), vp, i64]
    L.hvk_db_open.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
    L.hvk_db_close.argtypes = [vp]
    L.hvk_db_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(i64)]
    L.hvk_db_array.restype = vp
    L.hvk_db_array.argtypes = [vp, i32, ctypes.POINTER(i64)]
    L.hvk_db_upload.argtypes = [vp, vp, ctypes.POINTER(EncodeOut)]
    L.hvk_batch_create.argtypes = [i32, i32, i32, ctypes.POINTER(vp)]
    L.hvk_batch_destroy.argtypes = [vp]
    L.hvk_batch_config.argtypes = [vp, i32, i32, i32, i32]
    L.hvk_batch_get_config.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int)] * 5
    L.hvk_batch_classify.argtypes = [vp, vp, vp, vp, i32, ctypes.POINTER(Bundles), i32, vp, vp,
                                     vp, vp, vp]
    L.hvk_batch_autotune.argtypes = [vp, vp, vp, vp, i32, ctypes.POINTER(Bundles), i32, vp, vp,
                                     i32, ctypes.POINTER(ctypes.c_float), vp,
                                     ctypes.POINTER(ctypes.c_int), vp]
    L.hvk_batch_descend.argtypes = [vp, ctypes.POINTER(Tree), vp, vp, vp, i32, i32,
                                    ctypes.c_float, i32, vp, vp, vp, vp, vp]
    L.hvk_calib_mfma_i8.argtypes = [i32, i32, i32, i32, vp, ctypes.POINTER(ctypes.c_double), vp]
    _lib = L
    return L


def check(status: int, what: str):
    if status != HVK_OK:
        raise HvkError(status, what, lib())


# ----------------------------------------------------------------------------
# host helpers (no GPU)
# ----------------------------------------------------------------------------
def concat(seqs) -> tuple[bytes, np.ndarray]:
    """list[str|bytes] -> (concatenated bytes, int64 offsets[n+1])."""
    bs = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
    off = np.zeros(len(bs) + 1, np.int64)
    if bs:
        off[1:] = np.cumsum([len(b) for b in bs])
    return b"".join(bs), off


def pack(buf: bytes, offsets: np.ndarray):
    """2-bit codes + invalid-base mask (numpy uint32 arrays)."""
    L = lib()
    total = int(offsets[-1])
    words = int(L.hvk_pack_words(total))
    packed = np.zeros(words, np.uint32)
    invalid = np.zeros(words, np.uint32)
    cbuf = ctypes.create_string_buffer(buf, max(len(buf), 1))
    off = np.ascontiguousarray(offsets, np.int64)
    check(L.hvk_pack(cbuf, off.ctypes.data, len(off) - 1, packed.ctypes.data,
                     invalid.ctypes.data), "hvk_pack")
    return packed, invalid


@dataclass
class Workload:
    refs: bytes
    ref_offsets: np.ndarray
    reads: bytes
    read_offsets: np.ndarray
    truth: np.ndarray


def make_spec(**kw) -> GenSpec:
    d = dict(seed=0, n_ancestors=20, desc_per_ancestor=50, anc_len_min=1500, anc_len_max=1500,
             div_min=0.03, div_max=0.03, indel_rate=0.01, n_reads=500, read_len_min=250,
             read_len_max=250, read_len_log_uniform=0, err_min=0.01, err_max=0.01)
    d.update(kw)
    return GenSpec(**d)


def generate(spec: GenSpec) -> Workload:
    L = lib()
    a, b = ctypes.c_int64(), ctypes.c_int64()
    check(L.hvk_gen_size(ctypes.byref(spec), ctypes.byref(a), ctypes.byref(b)), "hvk_gen_size")
    n_refs = spec.n_ancestors * spec.desc_per_ancestor
    refs = ctypes.create_string_buffer(max(a.value, 1))
    reads = ctypes.create_string_buffer(max(b.value, 1))
    ro = np.zeros(n_refs + 1, np.int64)
    qo = np.zeros(spec.n_reads + 1, np.int64)
    truth = np.zeros(spec.n_reads, np.int32)
    check(L.hvk_gen_fill(ctypes.byref(spec), refs, ro.ctypes.data, reads, qo.ctypes.data,
                         truth.ctypes.data), "hvk_gen_fill")
    return Workload(refs.raw[: a.value], ro, reads.raw[: b.value], qo, truth)


# ----------------------------------------------------------------------------
# device side
# ----------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Context:
    """D, k and the device tables (hvk_ctx)."""

    def __init__(self, dim: int, kmer: int):
        L = lib()
        h = ctypes.c_void_p()
        check(L.hvk_ctx_create(dim, kmer, ctypes.byref(h)), "hvk_ctx_create")
        self.h = h
        d, k, w, ld = (ctypes.c_int() for _ in range(4))
        check(L.hvk_ctx_info(h, ctypes.byref(d), ctypes.byref(k), ctypes.byref(w), ctypes.byref(ld)),
              "hvk_ctx_info")
        self.dim, self.kmer, self.words, self.ld = d.value, k.value, w.value, ld.value

    def close(self):
        if self.h:
            lib().hvk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, param: int, value: int):
        check(lib().hvk_ctx_set_tuning(self.h, int(param), int(value)), "hvk_ctx_set_tuning")

    def set_timing(self, enable: bool):
        check(lib().hvk_ctx_set_timing(self.h, int(enable)), "hvk_ctx_set_timing")

    def last_timing(self):
        """(encode_ms, search_ms) of the main kernels of the last calls."""
        e, s = ctypes.c_float(), ctypes.c_float()
        check(lib().hvk_ctx_last_timing(self.h, ctypes.byref(e), ctypes.byref(s)),
              "hvk_ctx_last_timing")
        return e.value, s.value

    def base_hv(self, symbol: str) -> np.ndarray:
        out = np.zeros(self.words, np.uint64)
        check(lib().hvk_base_hv(self.h, symbol.encode(), out.ctypes.data), "hvk_base_hv")
        return out


@dataclass
class Encoded:
    n: int
    counts: object   # torch int16 [n][ld] or None
    lo: object       # torch int8 [n][ld] or None
    hi: object
    sqnorm: object   # torch int64 [n]
    maxabs: object   # torch int32 [n]
    nkmers: object   # torch int32 [n]
    clamp8: object = None  # torch int8 [n][ld] or None
    resid1: object = None  # torch int32 [n] or None


def _alloc(n, ld, counts, limbs, hi, clamp):
    torch = _torch()
    dev = "cuda"
    return Encoded(
        n,
        torch.empty((n, ld), dtype=torch.int16, device=dev) if counts else None,
        torch.empty((n, ld), dtype=torch.int8, device=dev) if limbs else None,
        torch.empty((n, ld), dtype=torch.int8, device=dev) if (limbs and hi) else None,
        torch.empty(n, dtype=torch.int64, device=dev),
        torch.empty(n, dtype=torch.int32, device=dev),
        torch.empty(n, dtype=torch.int32, device=dev),
        torch.empty((n, ld), dtype=torch.int8, device=dev) if clamp else None,
        torch.empty(n, dtype=torch.int32, device=dev) if clamp else None)


def _out_struct(e: Encoded, ld: int) -> EncodeOut:
    p = lambda t: None if t is None else t.data_ptr()
    return EncodeOut(p(e.counts), p(e.lo), p(e.hi), p(e.clamp8), ld, p(e.sqnorm), p(e.maxabs),
                     p(e.nkmers), p(e.resid1))


class DeviceSeqs:
    """Packed sequences resident on the device."""

    def __init__(self, buf: bytes, offsets: np.ndarray):
        torch = _torch()
        packed, invalid = pack(buf, offsets)
        self.n = len(offsets) - 1
        lens = np.diff(offsets)
        self.max_len = int(lens.max()) if self.n else 0
        self.packed = torch.from_numpy(packed.view(np.int32)).cuda()
        self.invalid = torch.from_numpy(invalid.view(np.int32)).cuda()
        self.host_offsets = np.ascontiguousarray(offsets, np.int64)
        self.offsets = torch.from_numpy(self.host_offsets).cuda()


def encode(ctx: Context, seqs: DeviceSeqs, counts=True, limbs=True, hi=True,
           clamp=False) -> Encoded:
    e = _alloc(seqs.n, ctx.ld, counts, limbs, hi, clamp)
    o = _out_struct(e, ctx.ld)
    check(lib().hvk_encode_ex(ctx.h, _ptr(seqs.packed), _ptr(seqs.invalid), _ptr(seqs.offsets),
                              seqs.n, seqs.max_len, ctypes.byref(o), _stream()), "hvk_encode_ex")
    return e


def encode_windows(ctx: Context, seqs: DeviceSeqs, starts: np.ndarray, window_len: int,
                   counts=True, limbs=True, hi=True, clamp=False) -> Encoded:
    """Bundles of the windows [starts[i], starts[i] + window_len) of the packed buffer."""
    torch = _torch()
    st = torch.from_numpy(np.ascontiguousarray(starts, np.int64)).cuda()
    e = _alloc(len(starts), ctx.ld, counts, limbs, hi, clamp)
    o = _out_struct(e, ctx.ld)
    check(lib().hvk_encode_windows(ctx.h, _ptr(seqs.packed), _ptr(seqs.invalid), _ptr(st),
                                   len(starts), window_len, ctypes.byref(o), _stream()),
          "hvk_encode_windows")
    torch.cuda.current_stream().synchronize()  # st is released on return
    return e


@dataclass
class TopK:
    idx: object     # torch int32 [n][k]
    score: object   # torch float32 [n][k]
    cos64: object   # torch float64 [n][k]
    n_rescored: object  # torch int32 [1]


def _topk_bufs(n, k):
    torch = _torch()
    return TopK(torch.empty((n, k), dtype=torch.int32, device="cuda"),
                torch.empty((n, k), dtype=torch.float32, device="cuda"),
                torch.empty((n, k), dtype=torch.float64, device="cuda"),
                torch.zeros(1, dtype=torch.int32, device="cuda"))


def search_plain(ctx: Context, q: Encoded, r: Encoded, k: int, q_maxabs=None, r_maxabs=None,
                 out: TopK | None = None) -> TopK:
    qm = int(q.maxabs.max().item()) if q_maxabs is None else q_maxabs
    rm = int(r.maxabs.max().item()) if r_maxabs is None else r_maxabs
    o = out or _topk_bufs(q.n, k)
    check(lib().hvk_search_plain(ctx.h, _ptr(q.counts), _ptr(q.sqnorm), q.n, qm, _ptr(r.counts),
                                 _ptr(r.sqnorm), r.n, rm, ctx.ld, k, _ptr(o.idx), _ptr(o.score),
                                 _ptr(o.cos64), _ptr(o.n_rescored), _stream()), "hvk_search_plain")
    return o


def search_mfma(ctx: Context, q: Encoded, r: Encoded, k: int, q_two_limbs=None, r_two_limbs=None,
                out: TopK | None = None) -> TopK:
    """q_two_limbs/r_two_limbs: None = decide from max |count| (> 127)."""
    if q_two_limbs is None:
        q_two_limbs = int(q.maxabs.max().item()) > 127
    if r_two_limbs is None:
        r_two_limbs = int(r.maxabs.max().item()) > 127
    o = out or _topk_bufs(q.n, k)
    check(lib().hvk_search_mfma(ctx.h, _ptr(q.lo), _ptr(q.hi) if q_two_limbs else None,
                                _ptr(q.sqnorm), q.n, _ptr(r.lo),
                                _ptr(r.hi) if r_two_limbs else None, _ptr(r.sqnorm), r.n, ctx.ld,
                                k, _ptr(o.idx), _ptr(o.score), _ptr(o.cos64), _ptr(o.n_rescored),
                                _stream()), "hvk_search_mfma")
    return o


def bundles(e: Encoded, two_limbs: bool | None = None) -> Bundles:
    p = lambda t: None if t is None else t.data_ptr()
    if two_limbs is None:
        two_limbs = int(e.maxabs.max().item()) > 127 if e.n else False
    return Bundles(p(e.clamp8), p(e.lo), p(e.hi) if two_limbs else None, p(e.sqnorm),
                   p(e.maxabs), p(e.resid1), e.n)


def search_mfma_bounded(ctx: Context, q: Encoded, r: Encoded, k: int,
                        out: TopK | None = None, q_two_limbs=None, r_two_limbs=None) -> TopK:
    """One-limb bounded MFMA form; q and r need clamp8/resid1
    (encode(..., clamp=True)) and their limbs."""
    o = out or _topk_bufs(q.n, k)
    bq, br = bundles(q, q_two_limbs), bundles(r, r_two_limbs)
    check(lib().hvk_search_mfma_bounded(ctx.h, ctypes.byref(bq), ctypes.byref(br), ctx.ld, k,
                                        _ptr(o.idx), _ptr(o.score), _ptr(o.cos64),
                                        _ptr(o.n_rescored), _stream()), "hvk_search_mfma_bounded")
    return o


def prepare_counts(ctx: Context, counts) -> Encoded:
    """Search operands (limbs, clamp, norms) from a torch int16 [n][ld] tensor."""
    n = counts.shape[0]
    e = _alloc(n, ctx.ld, counts=False, limbs=True, hi=True, clamp=True)
    e.counts = counts
    o = _out_struct(e, ctx.ld)
    o.counts = None
    check(lib().hvk_prepare_counts(ctx.h, _ptr(counts), n, ctypes.byref(o), _stream()),
          "hvk_prepare_counts")
    return e


# ----------------------------------------------------------------------------
# hierarchical descent
# ----------------------------------------------------------------------------
class Taxonomy:
    """Categories 0..n_cat-1 with candidate lists in CSR form (a category
    without children lists itself, as identify's get_children does), the
    level-0 candidates `roots`, and `ranks`: the category ids of each rank,
    top down.  `leaf_seqs`/`leaf_offsets`: the sequences of the last rank's
    categories, in ranks[-1] order."""

    def __init__(self, child_off, child_idx, roots, ranks, leaf_seqs=None, leaf_offsets=None):
        self.child_off = np.ascontiguousarray(child_off, np.int32)
        self.child_idx = np.ascontiguousarray(child_idx, np.int32)
        self.roots = np.ascontiguousarray(roots, np.int32)
        self.ranks = [np.ascontiguousarray(r, np.int32) for r in ranks]
        self.n_cat = len(self.child_off) - 1
        self.leaf_seqs = leaf_seqs
        self.leaf_offsets = leaf_offsets


def taxonomy_from_parents(parent):
    """CSR candidate lists from a parent array (parent[c] = -1 for the top
    rank); children keep id order, leaves list themselves, roots are the
    top rank in id order."""
    parent = np.asarray(parent, np.int64)
    n = len(parent)
    has_p = parent >= 0
    nkids = np.bincount(parent[has_p], minlength=n) if n else np.zeros(0, np.int64)
    leaf = np.nonzero(nkids == 0)[0]
    key = np.concatenate([parent[has_p], leaf])
    val = np.concatenate([np.nonzero(has_p)[0], leaf])
    o = np.lexsort((val, key))
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum(np.maximum(nkids, 1))
    depth = np.zeros(n, np.int64)
    for _ in range(n):
        nd = np.where(has_p, depth[np.maximum(parent, 0)] + 1, 0)
        if np.array_equal(nd, depth):
            break
        depth = nd
    ranks = [np.nonzero(depth == d)[0] for d in range(int(depth.max()) + 1)] if n else []
    return Taxonomy(off, val[o].astype(np.int32), np.nonzero(~has_p)[0].astype(np.int32), ranks)


def _mutate(rng, codes, rate):
    m = rng.random(codes.shape) < rate
    out = codes.copy()
    out[m] = (out[m] + rng.integers(1, 4, int(m.sum()), dtype=np.uint8)) % 4
    return out


def make_taxonomy(sizes=(40, 200, 800, 3000, 12000, 60000), seq_len=1500,
                  rates=(0.20, 0.10, 0.06, 0.04, 0.02, 0.01), n_reads=1000,
                  read_len=(150, 1600), err=0.01, seed=0, chunk=100_000):
    """Synthetic taxonomy: rank r has sizes[r] categories, category i of rank
    r has parent floor(i * sizes[r-1] / sizes[r]); each category's sequence
    is its parent's (a random root for the top rank) with a fraction
    rates[r] of substituted bases, so leaves are rates[-1] away from their
    parent centroid.  Reads: a uniform leaf, a log-uniform length in
    read_len clamped to the leaf length, a uniform start, `err`
    substitutions.  Returns (Taxonomy with leaf sequences, reads bytes,
    read offsets, leaf category of each read)."""
    rng = np.random.default_rng(seed)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    base = np.cumsum([0] + list(sizes))
    parent = np.full(base[-1], -1, np.int64)
    seq = rng.integers(0, 4, (1, seq_len), dtype=np.uint8)
    prev = None
    for r, n in enumerate(sizes):
        if r == 0:
            par_local = np.zeros(n, np.int64)
        else:
            par_local = (np.arange(n) * sizes[r - 1]) // n
            parent[base[r]:base[r + 1]] = base[r - 1] + par_local
        seq = _mutate(rng, seq[par_local], rates[r])
        prev = seq
    leaves = prev  # [sizes[-1]][seq_len] codes
    tax = taxonomy_from_parents(parent)
    tax.leaf_seqs = acgt[leaves].tobytes()
    tax.leaf_offsets = np.arange(sizes[-1] + 1, dtype=np.int64) * seq_len
    truth = rng.integers(0, sizes[-1], n_reads)
    lens = np.exp(rng.uniform(np.log(read_len[0]), np.log(read_len[1]), n_reads)).astype(np.int64)
    lens = np.minimum(lens, seq_len)
    starts = (rng.random(n_reads) * (seq_len - lens + 1)).astype(np.int64)
    offs = np.zeros(n_reads + 1, np.int64)
    offs[1:] = np.cumsum(lens)
    out = np.empty(int(offs[-1]), np.uint8)
    for a in range(0, n_reads, chunk):
        b = min(n_reads, a + chunk)
        ln = lens[a:b]
        pos = np.arange(int(offs[b] - offs[a])) - np.repeat(offs[a:b] - offs[a], ln)
        src = leaves[np.repeat(truth[a:b], ln), np.repeat(starts[a:b], ln) + pos]
        out[offs[a]:offs[b]] = acgt[_mutate(rng, src, err)]
    return tax, out.tobytes(), offs, (base[-2] + truth).astype(np.int32)


class TreeProto:
    """Device prototypes of a Taxonomy (int32 rows and their limbs)."""

    def __init__(self, ctx, tax: Taxonomy, proto, lo, hi, sqnorm, maxabs, row_maxabs):
        self.tax, self.proto, self.lo, self.hi, self.sqnorm = tax, proto, lo, hi, sqnorm
        self.row_maxabs = row_maxabs
        self.max_abs_limb = int(maxabs.max().item()) if maxabs.numel() else 0
        roots = _torch().from_numpy(tax.roots.astype(np.int64)).to(proto.device)
        self.max_abs_proto = (int(proto[roots].abs().max().item()) if len(tax.roots) else 0)
        self.ld = ctx.ld
        self.use_dense = True  # tests switch the level-0 matrix-core path off
        self.use_row_maxabs = True  # tests switch the per-row maxima off

    def struct(self) -> Tree:
        t = self.tax
        return Tree(t.n_cat, self.ld, self.lo.data_ptr(), self.hi.data_ptr(),
                    self.sqnorm.data_ptr(), self.max_abs_limb, t.child_off.ctypes.data,
                    t.child_idx.ctypes.data, t.roots.ctypes.data, len(t.roots),
                    self.max_abs_proto if self.use_dense else 0,
                    self.row_maxabs.data_ptr() if self.use_row_maxabs else None)


def build_prototypes(ctx: Context, tax: Taxonomy, leaf_counts, leaf_cat,
                     sum_parents=True) -> TreeProto:
    """Rows from encoded int16 counts (leaf_counts [n][ld] placed at
    categories leaf_cat), then, with sum_parents, every category with
    children set to the sum of its children rank by rank from the bottom,
    then limbs.  sum_parents=False keeps each placed row as its category's
    prototype."""
    torch = _torch()
    L = lib()
    n_cat, ld = tax.n_cat, ctx.ld
    proto = torch.zeros((n_cat, ld), dtype=torch.int32, device="cuda")
    lc = np.ascontiguousarray(leaf_cat, np.int32)
    check(L.hvk_tree_place(ctx.h, _ptr(leaf_counts), ld, len(lc), lc.ctypes.data, _ptr(proto), ld,
                           n_cat, _stream()), "hvk_tree_place")
    ovf = torch.zeros(1, dtype=torch.int32, device="cuda")
    span = np.diff(tax.child_off)
    self_list = (span == 1) & (tax.child_idx[tax.child_off[:-1]] == np.arange(n_cat))
    for rank in (reversed(tax.ranks[:-1]) if sum_parents else []):
        rank = np.ascontiguousarray(rank[~self_list[rank]], np.int32)
        check(L.hvk_tree_sum(ctx.h, _ptr(proto), ld, n_cat, tax.child_off.ctypes.data,
                             tax.child_idx.ctypes.data, rank.ctypes.data, len(rank), _ptr(ovf),
                             _stream()), "hvk_tree_sum")
    if int(ovf.item()):
        raise HvkError(HVK_ERR_RANGE, "prototype sums exceed int32")
    lo = torch.empty((n_cat, ld), dtype=torch.int16, device="cuda")
    hi = torch.empty_like(lo)
    sq = torch.empty(n_cat, dtype=torch.int64, device="cuda")
    mx = torch.empty(n_cat, dtype=torch.int32, device="cuda")
    mp = torch.empty(n_cat, dtype=torch.int32, device="cuda")
    check(L.hvk_tree_limbs(ctx.h, _ptr(proto), n_cat, ld, _ptr(lo), _ptr(hi), _ptr(sq), _ptr(mx),
                           _ptr(mp), _stream()), "hvk_tree_limbs")
    return TreeProto(ctx, tax, proto, lo, hi, sq, mx, mp)


class Descent:
    def __init__(self, path, score, length, exit_level):
        self.path, self.score, self.length, self.exit_level = path, score, length, exit_level


def descend(ctx: Context, tree: TreeProto, reads: Encoded, beam=5, threshold=0.0, levels=6,
            q_maxabs=None, out: Descent | None = None) -> Descent:
    """identify's beam over the tree's prototypes (hvk_descend); reads need
    int16 counts (encode(..., counts=True))."""
    torch = _torch()
    n = reads.n
    if out is None:
        out = Descent(torch.full((n, levels), -1, dtype=torch.int32, device="cuda"),
                      torch.zeros((n, levels), dtype=torch.float32, device="cuda"),
                      torch.zeros(n, dtype=torch.int32, device="cuda"),
                      torch.zeros(n, dtype=torch.int32, device="cuda"))
    if q_maxabs is None:
        q_maxabs = int(reads.maxabs.max().item()) if n else 0
    t = tree.struct()
    check(lib().hvk_descend(ctx.h, ctypes.byref(t), _ptr(reads.counts), _ptr(reads.sqnorm), n,
                            ctx.ld, q_maxabs, beam, threshold, levels, _ptr(out.path),
                            _ptr(out.score), _ptr(out.length), _ptr(out.exit_level), _stream()),
          "hvk_descend")
    return out


# ----------------------------------------------------------------------------
# batching form
# ----------------------------------------------------------------------------
class Batch:
    """Length-ordered, graph-replayed classification of large read batches
    (hvk_batch_*); owns a private context of the same D, k."""

    def __init__(self, dim: int, kmer: int, chunk: int = 262144):
        h = ctypes.c_void_p()
        check(lib().hvk_batch_create(dim, kmer, chunk, ctypes.byref(h)), "hvk_batch_create")
        self.h = h

    def close(self):
        if self.h:
            lib().hvk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self, form=BATCH_FORM_BOUNDED, splits=0, chunk=262144, shape=0):
        check(lib().hvk_batch_config(self.h, form, shape, splits, chunk), "hvk_batch_config")

    def get_config(self):
        v = [ctypes.c_int() for _ in range(5)]
        check(lib().hvk_batch_get_config(self.h, *[ctypes.byref(x) for x in v]),
              "hvk_batch_get_config")
        return dict(zip(("form", "shape", "splits", "chunk", "graphs"), (x.value for x in v)))

    def classify(self, seqs: DeviceSeqs, db: Bundles, k: int, out: TopK | None = None) -> TopK:
        o = out or _topk_bufs(seqs.n, k)
        check(lib().hvk_batch_classify(self.h, _ptr(seqs.packed), _ptr(seqs.invalid),
                                       seqs.host_offsets.ctypes.data, seqs.n, ctypes.byref(db), k,
                                       _ptr(o.idx), _ptr(o.score), _ptr(o.cos64),
                                       _ptr(o.n_rescored), _stream()), "hvk_batch_classify")
        return o

    def autotune(self, seqs: DeviceSeqs, db: Bundles, k: int, out: TopK, sample: int,
                 table: list | None = None) -> float:
        """Applies the fastest setting; returns its ms per million reads.  With
        `table` (a list), appends one dict per candidate timed."""
        ms = ctypes.c_float()
        rows = ctypes.c_int(64)
        tab = np.zeros((64, 5), np.float32)
        check(lib().hvk_batch_autotune(self.h, _ptr(seqs.packed), _ptr(seqs.invalid),
                                       seqs.host_offsets.ctypes.data, seqs.n, ctypes.byref(db), k,
                                       _ptr(out.idx), _ptr(out.score), sample, ctypes.byref(ms),
                                       tab.ctypes.data, ctypes.byref(rows), _stream()),
              "hvk_batch_autotune")
        if table is not None:
            for f, sh, sp, ch, t in tab[:rows.value]:
                table.append({"form": ["bounded", "two_limb"][int(f)], "shape": int(sh),
                              "splits": int(sp), "chunk": int(ch), "ms_per_mread": float(t)})
        return ms.value

    def descend(self, tree: "TreeProto", seqs: DeviceSeqs, beam=5, threshold=0.0, levels=6,
                out: "Descent | None" = None) -> "Descent":
        """hvk_batch_descend: encode + descent per length-ordered chunk, replayed
        from captured graphs; results in the reads' own rows."""
        torch = _torch()
        n = seqs.n
        if out is None:
            out = Descent(torch.full((n, levels), -1, dtype=torch.int32, device="cuda"),
                          torch.zeros((n, levels), dtype=torch.float32, device="cuda"),
                          torch.zeros(n, dtype=torch.int32, device="cuda"),
                          torch.zeros(n, dtype=torch.int32, device="cuda"))
        t = tree.struct()
        check(lib().hvk_batch_descend(self.h, ctypes.byref(t), _ptr(seqs.packed),
                                      _ptr(seqs.invalid), seqs.host_offsets.ctypes.data, n, beam,
                                      threshold, levels, _ptr(out.path), _ptr(out.score),
                                      _ptr(out.length), _ptr(out.exit_level), _stream()),
              "hvk_batch_descend")
        return out


# ----------------------------------------------------------------------------
# prototype database file
# ----------------------------------------------------------------------------
def db_write_host(path, dim, kmer, ld, arrays: dict, names: bytes | None = None):
    """Writes a database from host numpy arrays: keys counts (int16 [n][ld]),
    lo, hi, clamp8 (int8 [n][ld]), and sqnorm, maxabs, resid1, nkmers (all
    four or none)."""
    flags = 0
    o = EncodeOut()
    keep = []
    n = None
    for key, flag, dt in (("counts", DB_COUNTS, np.int16), ("lo", DB_LO, np.int8),
                          ("hi", DB_HI, np.int8), ("clamp8", DB_CLAMP8, np.int8)):
        if arrays.get(key) is not None:
            a = np.ascontiguousarray(arrays[key], dt)
            keep.append(a)
            setattr(o, key, a.ctypes.data)
            flags |= flag
            n = a.shape[0]
    if arrays.get("sqnorm") is not None:
        for key, dt in (("sqnorm", np.int64), ("maxabs", np.int32), ("resid1", np.int32),
                        ("nkmers", np.int32)):
            a = np.ascontiguousarray(arrays[key], dt)
            keep.append(a)
            setattr(o, key, a.ctypes.data)
            n = a.shape[0]
        flags |= DB_SUMS
    rows = [x for x in keep if x.ndim == 2]
    if any(x.shape[1] != rows[0].shape[1] for x in rows):
        raise ValueError("row arrays must share one stride")
    o.ld = rows[0].shape[1] if rows else ld  # source stride; the file's is ld
    nb = names or b""
    check(lib().hvk_db_write(path.encode(), dim, kmer, ld, n or 0, flags, ctypes.byref(o),
                             nb if nb else None, len(nb)), "hvk_db_write")


def db_save(path, ctx: Context, e: Encoded, names: bytes | None = None):
    """Writes an encoded database (device arrays of e) to `path`."""
    arrays = {k: (getattr(e, k).cpu().numpy() if getattr(e, k) is not None else None)
              for k in ("counts", "lo", "hi", "clamp8")}
    if e.resid1 is not None:
        arrays.update(sqnorm=e.sqnorm.cpu().numpy(), maxabs=e.maxabs.cpu().numpy(),
                      resid1=e.resid1.cpu().numpy(), nkmers=e.nkmers.cpu().numpy())
    db_write_host(path, ctx.dim, ctx.kmer, ctx.ld, arrays, names)


class DbFile:
    """A mapped database file (hvk_db_open)."""

    def __init__(self, path):
        h = ctypes.c_void_p()
        check(lib().hvk_db_open(path.encode(), ctypes.byref(h)), "hvk_db_open")
        self.h = h
        d, k, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ld, nl = ctypes.c_int64(), ctypes.c_int64()
        fl = ctypes.c_uint32()
        check(lib().hvk_db_info(h, ctypes.byref(d), ctypes.byref(k), ctypes.byref(ld),
                                ctypes.byref(n), ctypes.byref(fl), ctypes.byref(nl)), "hvk_db_info")
        self.dim, self.kmer, self.ld, self.n = d.value, k.value, ld.value, n.value
        self.flags, self.names_len = fl.value, nl.value

    def array(self, which, dtype, shape):
        b = ctypes.c_int64()
        p = lib().hvk_db_array(self.h, which, ctypes.byref(b))
        if not p:
            return None
        buf = (ctypes.c_char * b.value).from_address(p)
        return np.frombuffer(buf, dtype=dtype).reshape(shape)

    def names(self) -> bytes:
        b = ctypes.c_int64()
        p = lib().hvk_db_array(self.h, 64, ctypes.byref(b))
        return ctypes.string_at(p, b.value) if p else b""

    def load(self, ctx: Context) -> Encoded:
        """Device copy of the arrays (row stride ctx.ld)."""
        f = self.flags
        e = _alloc(self.n, ctx.ld, counts=bool(f & DB_COUNTS), limbs=bool(f & DB_LO),
                   hi=bool(f & DB_HI), clamp=bool(f & DB_CLAMP8))
        o = _out_struct(e, ctx.ld)
        check(lib().hvk_db_upload(self.h, ctx.h, ctypes.byref(o)), "hvk_db_upload")
        return e

    def close(self):
        if self.h:
            lib().hvk_db_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
