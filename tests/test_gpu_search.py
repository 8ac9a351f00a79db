"""GPU top-k search parity: plain (VALU) and MFMA forms vs the FP64
brute-force oracle (order: cosine desc, reference index asc)."""
import numpy as np
import pytest

import oracle_py as orc

pytestmark = pytest.mark.gpu

DIM, K = 10240, 12
COS_REL_TOL = 1e-6  # relative tolerance of the float32 score vs FP64


def encoded_from_counts(counts, ld):
    """Encoded bundle set built on the host from an int16 count matrix."""
    import hdc_amd as H
    import torch
    n, d = counts.shape
    c = np.zeros((n, ld), np.int16)
    c[:, :d] = counts
    v = c.astype(np.int32)
    lo = ((v + 128) % 256 - 128).astype(np.int8)
    hi = ((v - lo) // 256).astype(np.int8)
    cl = np.clip(v, -127, 127)
    sq = (c.astype(np.int64) ** 2).sum(1)
    mx = np.abs(v).max(1).astype(np.int32) if n else np.zeros(0, np.int32)
    res = np.abs(v - cl).sum(1).astype(np.int32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return H.Encoded(n, t(c), t(lo), t(hi), t(sq), t(mx), t(np.zeros(n, np.int32)),
                     t(cl.astype(np.int8)), t(res))


FORMS = ["plain", "mfma", "bounded"]


def run_form(form, ctx, q, r, k):
    import hdc_amd as H
    if form == "plain":
        return H.search_plain(ctx, q, r, k)
    if form == "mfma":
        return H.search_mfma(ctx, q, r, k)
    return H.search_mfma_bounded(ctx, q, r, k)


@pytest.fixture(scope="module")
def workload():
    import hdc_amd as H
    import torch
    spec = H.make_spec(n_ancestors=12, desc_per_ancestor=25, anc_len_min=1300, anc_len_max=1600,
                       div_min=0.02, div_max=0.08, n_reads=157, read_len_min=150,
                       read_len_max=300, err_min=0.005, err_max=0.02)
    w = H.generate(spec)
    ctx = H.Context(DIM, K)
    refs = H.encode(ctx, H.DeviceSeqs(w.refs, w.ref_offsets), clamp=True)
    reads = H.encode(ctx, H.DeviceSeqs(w.reads, w.read_offsets), clamp=True)
    torch.cuda.synchronize()
    rc = refs.counts.cpu().numpy()[:, :DIM]
    qc = reads.counts.cpu().numpy()[:, :DIM]
    return ctx, refs, reads, rc, qc, w


def _check(res, ref_idx, ref_cos, k):
    import torch
    torch.cuda.synchronize()
    idx = res.idx.cpu().numpy()
    sc = res.score.cpu().numpy().astype(np.float64)
    c64 = res.cos64.cpu().numpy()
    assert np.array_equal(idx, ref_idx[:, :k])
    ok = ref_idx[:, :k] >= 0
    assert np.allclose(c64[ok], ref_cos[:, :k][ok], rtol=1e-12, atol=1e-15)
    err = np.abs(sc[ok] - ref_cos[:, :k][ok]) / np.maximum(np.abs(ref_cos[:, :k][ok]), 1e-30)
    assert err.max(initial=0.0) <= COS_REL_TOL


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("k", [1, 5, 10])
def test_topk_matches_oracle(workload, form, k):
    import hdc_amd as H
    ctx, refs, reads, rc, qc, w = workload
    print(f"search form={form} k={k} nq={reads.n} nr={refs.n}", flush=True)
    ref_idx, ref_cos = orc.topk16(qc, rc, k)
    res = run_form(form, ctx, reads, refs, k)
    _check(res, ref_idx, ref_cos, k)


@pytest.mark.parametrize("form", FORMS)
def test_reads_find_their_source(workload, form):
    """Top-1 of a read sampled from reference r is r (or an equal-score tie)."""
    import hdc_amd as H
    ctx, refs, reads, rc, qc, w = workload
    res = run_form(form, ctx, reads, refs, 5)
    idx = res.idx.cpu().numpy()
    hit = (idx[:, 0] == w.truth).mean()
    assert hit > 0.9


@pytest.mark.parametrize("form", FORMS)
def test_exact_ties_duplicate_references(workload, form):
    """Duplicated references tie exactly; order must be index ascending and
    the near-tie margin check must route the queries to the exact re-scan."""
    import hdc_amd as H
    ctx, refs, reads, rc, qc, w = workload
    # pairs of duplicates: ties inside the candidate list, no re-scan needed
    dup = np.concatenate([rc[:40], rc[:40], rc[5:6], rc[5:6], rc[5:6], rc[40:90]])
    db = encoded_from_counts(dup, ctx.ld)
    q = encoded_from_counts(qc[:33], ctx.ld)
    ref_idx, ref_cos = orc.topk16(qc[:33], dup, 10)
    res = run_form(form, ctx, q, db, 10)
    _check(res, ref_idx, ref_cos, 10)
    # 41 copies tie at c = 1 for queries equal to reference 5: the tie group
    # outgrows every candidate list (16 / 32), so those queries go through the
    # exact passes down to the brute-force re-scan
    big = np.concatenate([rc[:30], np.repeat(rc[5:6], 40, axis=0), rc[30:60]])
    qn = np.concatenate([rc[5:6], qc[:7], rc[5:6]])
    ref_idx, ref_cos = orc.topk16(qn, big, 10)
    res = run_form(form, ctx, encoded_from_counts(qn, ctx.ld), encoded_from_counts(big, ctx.ld), 10)
    _check(res, ref_idx, ref_cos, 10)
    assert list(ref_idx[0]) == [5] + list(range(30, 39))
    assert int(res.n_rescored.cpu().item()) >= 2


@pytest.mark.parametrize("form", FORMS)
def test_zero_norm_rows(workload, form):
    import hdc_amd as H
    ctx, refs, reads, rc, qc, w = workload
    db_np = rc[:70].copy()
    db_np[3] = 0
    db_np[64] = 0
    q_np = qc[:5].copy()
    q_np[2] = 0
    db = encoded_from_counts(db_np, ctx.ld)
    q = encoded_from_counts(q_np, ctx.ld)
    ref_idx, ref_cos = orc.topk16(q_np, db_np, 5)
    res = run_form(form, ctx, q, db, 5)
    _check(res, ref_idx, ref_cos, 5)
    assert list(res.idx.cpu().numpy()[2]) == [0, 1, 2, 3, 4]


@pytest.mark.parametrize("form", FORMS)
def test_k_larger_than_db(workload, form):
    import hdc_amd as H
    ctx, refs, reads, rc, qc, w = workload
    db = encoded_from_counts(rc[:3], ctx.ld)
    q = encoded_from_counts(qc[:4], ctx.ld)
    res = run_form(form, ctx, q, db, 5)
    idx = res.idx.cpu().numpy()
    ref_idx, _ = orc.topk16(qc[:4], rc[:3], 5)
    assert np.array_equal(idx, ref_idx)
    assert (idx[:, 3:] == -1).all()


def test_two_limb_queries(workload):
    """Queries with |count| > 127 take the two-limb MFMA path."""
    import hdc_amd as H
    import torch
    ctx, refs, reads, rc, qc, w = workload
    seqs = [w.refs[w.ref_offsets[i]:w.ref_offsets[i + 1]].decode() for i in (0, 30, 77)]
    seqs += ["ACGTTGCA" * 150, "A" * 800]
    buf, off = H.concat(seqs)
    q = H.encode(ctx, H.DeviceSeqs(buf, off), clamp=True)
    torch.cuda.synchronize()
    assert int(q.maxabs.max().item()) > 127
    qn = q.counts.cpu().numpy()[:, :DIM]
    ref_idx, ref_cos = orc.topk16(qn, rc, 5)
    for fn in (H.search_mfma, H.search_mfma_bounded):
        res = fn(ctx, q, refs, 5)
        _check(res, ref_idx, ref_cos, 5)
        assert list(res.idx.cpu().numpy()[0][:1]) == [0]


def test_bipolar_single_limb():
    """Binarised bipolar HVs (+1/-1) as one int8 limb on both sides."""
    import hdc_amd as H
    rng = np.random.default_rng(3)
    ctx = H.Context(DIM, K)
    db_np = np.where(rng.random((300, DIM)) < 0.5, 1, -1).astype(np.int16)
    q_np = db_np[rng.integers(0, 300, 64)].copy()
    flip = rng.random(q_np.shape) < 0.2
    q_np[flip] *= -1
    db = encoded_from_counts(db_np, ctx.ld)
    q = encoded_from_counts(q_np, ctx.ld)
    ref_idx, ref_cos = orc.topk16(q_np, db_np, 5)
    res = H.search_mfma(ctx, q, db, 5, q_two_limbs=False, r_two_limbs=False)
    _check(res, ref_idx, ref_cos, 5)


def test_bounded_second_pass_close_siblings():
    """Close siblings (0.2-1% divergence, 100 per ancestor): the bounded
    form's 8-entry lists for top-5 fail the margin check for some reads, and
    the exact second pass settles them.  Its results equal the two-limb
    form's over all reads and the FP64 oracle's on a sample."""
    import hdc_amd as H
    import torch
    spec = H.make_spec(seed=5, n_ancestors=8, desc_per_ancestor=100, anc_len_min=1300,
                       anc_len_max=1600, div_min=0.002, div_max=0.01, n_reads=3000,
                       read_len_min=150, read_len_max=300, err_min=0.005, err_max=0.02)
    w = H.generate(spec)
    ctx = H.Context(DIM, K)
    db = H.encode(ctx, H.DeviceSeqs(w.refs, w.ref_offsets), clamp=True)
    q = H.encode(ctx, H.DeviceSeqs(w.reads, w.read_offsets), clamp=True)
    b = H.search_mfma_bounded(ctx, q, db, 5)
    t = H.search_mfma(ctx, q, db, 5)
    torch.cuda.synchronize()
    n_res = int(b.n_rescored.item())
    print(f"bounded second pass: {n_res} of {q.n} reads re-scored", flush=True)
    assert n_res > 0
    assert torch.equal(b.idx, t.idx)
    assert torch.equal(b.cos64, t.cos64)
    rc = db.counts.cpu().numpy()[:, :DIM]
    qc = q.counts.cpu().numpy()[:256, :DIM]
    ref_idx, ref_cos = orc.topk16(qc, rc, 5)
    res = H.TopK(b.idx[:256], b.score[:256], b.cos64[:256], b.n_rescored)
    _check(res, ref_idx, ref_cos, 5)


# ---- k past the candidate lists: key matrix + radix select (any k) ----

@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("k", [30, 64, 300, 357])
def test_large_k_matches_oracle(workload, form, k):
    """k above the list length (29) up to and past the number of references
    (300 here): identical indices and FP64 values, -1 past the references."""
    ctx, refs, reads, rc, qc, w = workload
    print(f"large-k form={form} k={k} nq={reads.n} nr={refs.n}", flush=True)
    ref_idx, ref_cos = orc.topk16(qc, rc, k)
    res = run_form(form, ctx, reads, refs, k)
    _check(res, ref_idx, ref_cos, k)
    assert int(res.n_rescored.cpu().item()) == 0
    if k > refs.n:
        assert (res.idx.cpu().numpy()[:, refs.n:] == -1).all()


@pytest.fixture(scope="module")
def big_db():
    """1,100 references (44 ancestors x 25) and 48 reads, for k = 1024."""
    import hdc_amd as H
    import torch
    spec = H.make_spec(seed=11, n_ancestors=44, desc_per_ancestor=25, anc_len_min=1300,
                       anc_len_max=1600, div_min=0.02, div_max=0.08, n_reads=48,
                       read_len_min=150, read_len_max=300, err_min=0.005, err_max=0.02)
    w = H.generate(spec)
    ctx = H.Context(DIM, K)
    refs = H.encode(ctx, H.DeviceSeqs(w.refs, w.ref_offsets), clamp=True)
    reads = H.encode(ctx, H.DeviceSeqs(w.reads, w.read_offsets), clamp=True)
    torch.cuda.synchronize()
    return ctx, refs, reads, refs.counts.cpu().numpy()[:, :DIM], reads.counts.cpu().numpy()[:, :DIM]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("k", [1024, 1100])
def test_large_k_1024_and_all(big_db, form, k):
    ctx, refs, reads, rc, qc = big_db
    print(f"large-k form={form} k={k} nq={reads.n} nr={refs.n}", flush=True)
    ref_idx, ref_cos = orc.topk16(qc, rc, k)
    res = run_form(form, ctx, reads, refs, k)
    _check(res, ref_idx, ref_cos, k)


@pytest.mark.parametrize("form", FORMS)
def test_large_k_ties_and_zero_norms(workload, form):
    """40 identical references tie exactly (index order), a zero row scores
    c = 0, and a zero query ranks every reference at c = 0 (index order)."""
    ctx, refs, reads, rc, qc, w = workload
    big = np.concatenate([rc[:30], np.repeat(rc[5:6], 40, axis=0), rc[30:60]])
    big[77] = 0
    qn = np.concatenate([rc[5:6], qc[:7], np.zeros((1, DIM), np.int16)])
    for k in (45, 100):
        ref_idx, ref_cos = orc.topk16(qn, big, k)
        res = run_form(form, ctx, encoded_from_counts(qn, ctx.ld), encoded_from_counts(big, ctx.ld),
                       k)
        _check(res, ref_idx, ref_cos, k)
        assert list(res.idx.cpu().numpy()[8]) == list(range(k))
        assert list(ref_idx[0][:41]) == [5] + list(range(30, 70))


@pytest.mark.parametrize("form", FORMS)
def test_large_k_chunked(workload, form):
    """A 256 KiB selection budget splits the 157 queries into several chunks."""
    import hdc_amd as H
    ctx, refs, reads, rc, qc, w = workload
    ref_idx, ref_cos = orc.topk16(qc, rc, 64)
    ctx.set_tuning(H.TUNE_SELECT_KIB, 256)
    try:
        res = run_form(form, ctx, reads, refs, 64)
        _check(res, ref_idx, ref_cos, 64)
    finally:
        ctx.set_tuning(H.TUNE_SELECT_KIB, 0)


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5])
def test_bounded_tile_shapes(workload, shape):
    """Every tile shape of the bounded form (HVK_TUNE_SHAPE) gives the oracle's
    top-5, with exact ties and the second pass (41 copies) included."""
    import hdc_amd as H
    ctx, refs, reads, rc, qc, w = workload
    ctx.set_tuning(H.TUNE_SHAPE, shape)
    try:
        print(f"bounded shape={shape}", flush=True)
        ref_idx, ref_cos = orc.topk16(qc, rc, 5)
        _check(H.search_mfma_bounded(ctx, reads, refs, 5), ref_idx, ref_cos, 5)
        big = np.concatenate([rc[:30], np.repeat(rc[5:6], 40, axis=0), rc[30:60]])
        qn = np.concatenate([rc[5:6], qc[:7], rc[5:6]])
        ref_idx, ref_cos = orc.topk16(qn, big, 5)
        _check(H.search_mfma_bounded(ctx, encoded_from_counts(qn, ctx.ld),
                                     encoded_from_counts(big, ctx.ld), 5), ref_idx, ref_cos, 5)
    finally:
        ctx.set_tuning(H.TUNE_SHAPE, 0)


@pytest.fixture(scope="module")
def wide_workload():
    """2,000 references (8 reference tiles of 256) and 300 reads, so that 8
    slices can be grouped per XCD in several ways."""
    import hdc_amd as H
    import torch
    spec = H.make_spec(n_ancestors=40, desc_per_ancestor=50, anc_len_min=1300, anc_len_max=1600,
                       div_min=0.02, div_max=0.08, n_reads=300, read_len_min=150,
                       read_len_max=300, err_min=0.005, err_max=0.02)
    w = H.generate(spec)
    ctx = H.Context(DIM, K)
    refs = H.encode(ctx, H.DeviceSeqs(w.refs, w.ref_offsets), clamp=True)
    reads = H.encode(ctx, H.DeviceSeqs(w.reads, w.read_offsets), clamp=True)
    torch.cuda.synchronize()
    rc = refs.counts.cpu().numpy()[:, :DIM]
    qc = reads.counts.cpu().numpy()[:, :DIM]
    return ctx, refs, reads, rc, qc


@pytest.mark.parametrize("form", ["bounded", "two_limb"])
def test_xcd_slice_groups(wide_workload, form):
    """8 reference slices grouped 1, 3 (a partial last group), 8 or
    automatically per XCD (HVK_TUNE_GROUP_SLICES) give the oracle's top-5."""
    import hdc_amd as H
    ctx, refs, reads, rc, qc = wide_workload
    assert refs.n == 2000
    ref_idx, ref_cos = orc.topk16(qc, rc, 5)
    ctx.set_tuning(H.TUNE_SPLITS, 8)
    try:
        for sg in (1, 3, 8, 0):
            print(f"{form} slices=8 group_slices={sg}", flush=True)
            ctx.set_tuning(H.TUNE_GROUP_SLICES, sg)
            if form == "bounded":
                res = H.search_mfma_bounded(ctx, reads, refs, 5)
            else:
                res = H.search_mfma(ctx, reads, refs, 5, q_two_limbs=True, r_two_limbs=True)
            _check(res, ref_idx, ref_cos, 5)
    finally:
        ctx.set_tuning(H.TUNE_GROUP_SLICES, 0)
        ctx.set_tuning(H.TUNE_SPLITS, 0)
