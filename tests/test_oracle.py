"""Pins the C oracle (oracle/hdc_oracle.c) before it is trusted as the checker.

Each oracle function is checked against something other than its own
formula: the C++ standard's known answer, the reference's own CPU sources
(oracle/_ref, or the golden file they wrote), bit-level brute force written
independently here in numpy, and closed forms / invariants.
"""
import os

import numpy as np
import pytest

import oracle_py as orc
import seqgen

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_d10000_k5.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def bits_of(hv, D):
    """uint64 words -> bool array of the D live bits (bit d = word d/64 bit d%64)."""
    b = np.unpackbits(hv.view(np.uint8), bitorder="little").astype(bool)
    return b[:D]


def ring_bits(hv, D):
    W = (D + 63) // 64
    return np.unpackbits(hv.view(np.uint8), bitorder="little").astype(bool)[: 64 * W]


RC = str.maketrans("ACGT", "TGCA")


def canonical(s):
    rc = s.translate(RC)[::-1]
    return min(s, rc)


def kmer_bits_closed_form(D, k, kmer, base_bits):
    """HV(kmer)[d] = NOT XOR_i base[c_i][(d + i) mod R], R = 64 W (zero pad)."""
    R = 64 * ((D + 63) // 64)
    c = canonical(kmer)
    acc = np.zeros(D, bool)
    d = np.arange(D)
    for i, ch in enumerate(c):
        acc ^= base_bits[ch][(d + i) % R]
    return ~acc


def base_ring(D):
    return {ch: ring_bits(orc.base_hv(D, ch), D) for ch in "ACGT"}


# -- mt19937_64 ---------------------------------------------------------------
def test_mt19937_64_known_answer():
    # C++ [rand.predef]: the 10000th output of a default-constructed
    # mt19937_64 (seed 5489) is 9981545732273789042.
    assert int(orc.mt19937_64(5489, 10000)[-1]) == 9981545732273789042


def test_base_hv_matches_reference(gold):
    for i, ch in enumerate("ACGTN"):
        assert np.array_equal(orc.base_hv(10000, ch), gold["base_hv"][i])


def test_base_hv_tail_masked():
    for D in (10000, 100, 8192):
        for ch in "ACGTN":
            h = orc.base_hv(D, ch)
            r = ring_bits(h, D)
            assert not r[D:].any()
    assert bits_of(orc.base_hv(10000, "N"), 10000).all()


# -- mult ----------------------------------------------------------------------
@pytest.mark.parametrize("D", [100, 192, 10000, 10240])
@pytest.mark.parametrize("s", [0, 1, 5, 63, 64, 65, 130])
def test_mult_bit_level(D, s):
    W = (D + 63) // 64
    R = 64 * W
    rng = np.random.default_rng(D + s)
    l = rng.integers(0, 2**63, W, dtype=np.uint64)
    r = rng.integers(0, 2**63, W, dtype=np.uint64)
    out = np.zeros(W, np.uint64)
    orc.oracle().orc_mult(D, l.ctypes.data, r.ctypes.data, out.ctypes.data, s)
    lb, rb = ring_bits(l, D), ring_bits(r, D)
    want = lb ^ rb[(np.arange(R) + s) % R]
    want[D:] = False
    assert np.array_equal(ring_bits(out, D), want)


def test_mult_matches_reference():
    R = orc.reference()
    if R is None:
        pytest.skip("reference build absent (oracle/_ref)")
    D = 10000
    W = (D + 63) // 64
    rng = np.random.default_rng(5)
    for s in (0, 3, 64, 100, 9999):
        l = rng.integers(0, 2**63, W, dtype=np.uint64)
        r = rng.integers(0, 2**63, W, dtype=np.uint64)
        a = np.zeros(W, np.uint64)
        b = np.zeros(W, np.uint64)
        orc.oracle().orc_mult(D, l.ctypes.data, r.ctypes.data, a.ctypes.data, s)
        R.ref_mult(l.ctypes.data, r.ctypes.data, b.ctypes.data, s)
        assert np.array_equal(a, b)


# -- k-mer HVs -------------------------------------------------------------------
def test_kmer_hv_matches_reference(gold):
    for i, km in enumerate(gold["kmers"]):
        assert np.array_equal(orc.kmer_hv(10000, 5, str(km)), gold["kmer_hv"][i])


@pytest.mark.parametrize("D,k", [(10000, 5), (8192, 10), (10240, 12)])
def test_kmer_hv_closed_form_and_revcomp(D, k):
    rng = np.random.default_rng(k)
    br = base_ring(D)
    kms = [seqgen.random_seq(rng, k) for _ in range(6)]
    if k % 2 == 0:
        half = seqgen.random_seq(rng, k // 2)
        kms.append(half + half.translate(RC)[::-1])  # palindrome
    for km in kms:
        h = orc.kmer_hv(D, k, km)
        assert np.array_equal(bits_of(h, D), kmer_bits_closed_form(D, k, km, br))
        assert np.array_equal(h, orc.kmer_hv(D, k, km.translate(RC)[::-1]))
    assert orc.kmer_hv(D, k, "N" + "A" * (k - 1)) is None
    assert orc.kmer_hv(D, k, "a" * k) is None


# -- compute ---------------------------------------------------------------------
def test_compute_matches_reference(gold):
    for i, s in enumerate(gold["seqs"]):
        c, _ = orc.compute(10000, 5, str(s))
        assert np.array_equal(c, gold["counts"][i]), i


def test_compute_matches_live_reference():
    R = orc.reference()
    if R is None:
        pytest.skip("reference build absent (oracle/_ref)")
    for s in seqgen.mixed_set(9, 12, 1, 600):
        c, _ = orc.compute(10000, 5, s)
        r = np.zeros(10000, np.int16)
        b = s.encode()
        R.ref_compute(b, len(b), r.ctypes.data)
        assert np.array_equal(c, r)


@pytest.mark.parametrize("D,k", [(8192, 10), (10240, 12)])
def test_compute_bit_level_sum(D, k):
    """Independent numpy bundle of closed-form k-mer HVs at the bench configs."""
    br = base_ring(D)
    for s in seqgen.mixed_set(100 + k, 4, 20, 90):
        c, n = orc.compute(D, k, s)
        want = np.zeros(D, np.int64)
        cnt = 0
        for j in range(len(s) - k + 1):
            km = s[j:j + k]
            if any(ch not in "ACGT" for ch in km):
                continue
            want += np.where(kmer_bits_closed_form(D, k, km, br), 1, -1)
            cnt += 1
        assert n == cnt
        assert np.array_equal(c.astype(np.int64), want)


@pytest.mark.parametrize("D,k", [(10240, 12)])
def test_compute_invariants(D, k):
    for s in seqgen.mixed_set(5, 6, 100, 400):
        c, n = orc.compute(D, k, s)
        c32, n32 = orc.compute32(D, k, s)
        assert n == n32 and np.array_equal(c.astype(np.int32), c32)
        assert (np.abs(c32) <= n).all()
        assert ((c32 - n) % 2 == 0).all()


# -- superpose / replace ----------------------------------------------------------
def test_superpose_replace_identities():
    D = 10000
    W = (D + 63) // 64
    rng = np.random.default_rng(11)
    hvs = rng.integers(0, 2**63, (16, W), dtype=np.uint64)
    hvs[:, -1] &= np.uint64((1 << (D - 64 * (W - 1))) - 1)
    ptrs = (np.ctypeslib.ctypes.c_void_p * 16)(*[hvs[i].ctypes.data for i in range(16)])
    O = orc.oracle()
    acc = np.zeros(D, np.int16)
    O.orc_superpose_batch(D, acc.ctypes.data, ptrs, 16)
    one = np.zeros(D, np.int16)
    for i in range(16):
        p1 = (np.ctypeslib.ctypes.c_void_p * 1)(hvs[i].ctypes.data)
        O.orc_superpose_batch(D, one.ctypes.data, p1, 1)
    assert np.array_equal(acc, one)
    ones = np.zeros(W, np.uint64)
    O.orc_empty_hv(D, ones.ctypes.data)
    po = (np.ctypeslib.ctypes.c_void_p * 5)(*[ones.ctypes.data] * 5)
    z = np.zeros(D, np.int16)
    O.orc_superpose_batch(D, z.ctypes.data, po, 5)
    assert (z == 5).all()
    # replace(new, old) == superpose(new) - superpose(old)
    a = np.zeros(D, np.int16)
    new = (np.ctypeslib.ctypes.c_void_p * 8)(*[hvs[i].ctypes.data for i in range(8)])
    old = (np.ctypeslib.ctypes.c_void_p * 8)(*[hvs[i].ctypes.data for i in range(8, 16)])
    O.orc_replace_batch(D, a.ctypes.data, old, new, 8)
    sn = np.zeros(D, np.int16)
    so = np.zeros(D, np.int16)
    O.orc_superpose_batch(D, sn.ctypes.data, new, 8)
    O.orc_superpose_batch(D, so.ctypes.data, old, 8)
    assert np.array_equal(a, sn - so)
    R = orc.reference()
    if R is not None:
        b = np.zeros(D, np.int16)
        R.ref_superpose_batch(b.ctypes.data, hvs.ctypes.data, 16)
        assert np.array_equal(acc, b)
        c = np.zeros(D, np.int16)
        R.ref_replace_batch(c.ctypes.data, np.ascontiguousarray(hvs[8:]).ctypes.data,
                            np.ascontiguousarray(hvs[:8]).ctypes.data, 8)
        assert np.array_equal(a, c)


# -- cosine ------------------------------------------------------------------------
def test_cosine_matches_reference(gold):
    C = gold["counts"]
    n = len(C)
    for i in range(n):
        for j in range(n):
            assert orc.cosine(10000, C[i], C[j]) == gold["cosine"][i, j]


def test_cosine_closed_form():
    rng = np.random.default_rng(1)
    a = rng.integers(-300, 300, 10240).astype(np.int16)
    b = rng.integers(-300, 300, 10240).astype(np.int16)
    d = float(np.dot(a.astype(np.int64), b.astype(np.int64)))
    na, nb = float((a.astype(np.int64) ** 2).sum()), float((b.astype(np.int64) ** 2).sum())
    c = d / (np.sqrt(na) * np.sqrt(nb))
    assert orc.cosine(10240, a, b) == np.float32(2.0 * (c - 0.5))
    assert orc.cosine(10240, a, a) == np.float32(1.0)
    assert orc.cosine(10240, np.zeros(10240, np.int16), b) == np.float32(-1.0)
    assert orc.cosine(10240, a, b, norm_mode=0) == np.float32(d)


# -- precompute / similarity ------------------------------------------------------
def test_precompute_windows_match_reference(gold):
    rows = gold["win_rows"]
    o = 0
    for r, n in zip(gold["win_refs"], gold["win_counts"]):
        w = orc.precompute_windows(10000, 5, 125, 10, str(r))
        assert len(w) == n
        assert np.array_equal(w, rows[o:o + n])
        o += n


def test_similarity_matches_reference(gold):
    for i, q in enumerate(gold["queries"]):
        for j, r in enumerate(gold["win_refs"]):
            s, p = orc.precompute_similarity(10000, 5, 125, 10, str(q), str(r))
            assert s == gold["sim"][i, j] and p == gold["pos"][i, j]


@pytest.mark.parametrize("D,k,win,step", [(10000, 5, 125, 10), (2048, 7, 60, 6)])
def test_precompute_drift_closed_form(D, k, win, step):
    """Without cache hits, window w = compute(exact window) + bip(H_{win-k})
    - bip(H_{step*w + win - k}) (the reference's sliding-update offset)."""
    rng = np.random.default_rng(D)
    r = seqgen.random_seq(rng, 400)
    wins = orc.precompute_windows(D, k, win, step, r)
    assert len(wins) == len(range(0, len(r) - win, step))
    br = base_ring(D)
    bip = lambda km: np.where(kmer_bits_closed_form(D, k, km, br), 1, -1)
    for w in range(len(wins)):
        exact, _ = orc.compute32(D, k, r[step * w: step * w + win])
        want = exact.astype(np.int64)
        if w >= 1:
            want += bip(r[win - k: win]) - bip(r[step * w + win - k: step * w + win])
        assert np.array_equal(wins[w].astype(np.int64), want), w


# -- top-k and the CPU baseline path -------------------------------------------------
def test_topk_brute_force():
    rng = np.random.default_rng(4)
    db = rng.integers(-50, 50, (300, 256)).astype(np.int16)
    db[17] = db[3]          # exact tie: lower index first
    db[40] = 0              # zero norm: cosine 0
    q = rng.integers(-50, 50, (9, 256)).astype(np.int16)
    q[2] = db[3]
    idx, cos = orc.topk16(q, db, 6)
    d = q.astype(np.int64) @ db.T.astype(np.int64)
    nq = np.sqrt((q.astype(np.int64) ** 2).sum(1).astype(float))
    nd = np.sqrt((db.astype(np.int64) ** 2).sum(1).astype(float))
    with np.errstate(invalid="ignore", divide="ignore"):
        c = d / (nq[:, None] * nd[None, :])
    c[:, nd == 0] = 0.0
    for i in range(len(q)):
        order = np.lexsort((np.arange(300), -c[i]))[:6]
        assert np.array_equal(idx[i], order)
        assert np.allclose(cos[i], c[i][order], rtol=0, atol=1e-15)
    assert list(idx[2][:2]) == [3, 17]


def test_cpu_baseline_equals_compute_plus_topk():
    import ctypes
    D, k = 2048, 7
    rng = np.random.default_rng(8)
    refs = [seqgen.random_seq(rng, 300) for _ in range(40)]
    db = np.stack([orc.compute(D, k, s)[0] for s in refs])
    reads = [seqgen.mutate(rng, refs[i][50:200], 0.02) for i in range(0, 40, 4)]
    buf = "".join(reads).encode()
    off = np.zeros(len(reads) + 1, np.int64)
    off[1:] = np.cumsum([len(x) for x in reads])
    oi = np.zeros((len(reads), 5), np.int32)
    oc = np.zeros((len(reads), 5), np.float64)
    dbn = (db.astype(np.int64) ** 2).sum(1)
    orc.oracle().orc_classify_reads(D, k, buf, off.ctypes.data, len(reads), db.ctypes.data,
                                    dbn.ctypes.data, len(refs), 5, oi.ctypes.data,
                                    oc.ctypes.data, 3)
    q = np.stack([orc.compute(D, k, s)[0] for s in reads])
    idx, cos = orc.topk16(q, db, 5)
    assert np.array_equal(oi, idx) and np.array_equal(oc, cos)
    assert list(idx[:, 0]) == list(range(0, 40, 4))
