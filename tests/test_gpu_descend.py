"""GPU hierarchical descent (hvk_descend) against the reference and the oracle.

* tests/golden/tax_short: the reference's own identify output on a taxonomy
  whose categories each hold one 100-nt sequence (D=10000, k=5, beam 5),
  with the reference's candidate order; the result lines must be identical,
  ties included.
* Synthetic taxonomies (the 7-rank generator at small sizes, D=10240, k=12):
  prototypes built on the GPU (leaf bundles placed, parents summed) equal the
  oracle's, and descent paths / per-rank scores / lengths / exit levels equal
  orc_descend's bit for bit, for several beams and thresholds, with tied
  siblings (more than 16 candidates a level, so std::sort's introsort order
  decides), zero-norm and short reads, and the narrower partial-sum chunks.
"""
import numpy as np
import pytest

import oracle_py as orc
from test_descend_oracle import (D_REF, K_REF, candidate_order, expected_lines, format_lines,
                                 load_tax_short)

pytestmark = pytest.mark.gpu


def _dev_seqs(H, seqs):
    buf, off = H.concat(seqs)
    return H.DeviceSeqs(buf, off)


def test_descend_matches_reference_identify():
    import hdc_amd as H
    import torch
    names, parent, seqs, reads = load_tax_short()
    off, idx, roots = candidate_order(names)
    tax = H.Taxonomy(off, idx, roots, H.taxonomy_from_parents(parent).ranks)
    ctx = H.Context(D_REF, K_REF)
    cat_enc = H.encode(ctx, _dev_seqs(H, seqs), counts=True, limbs=False, hi=False)
    tree = H.build_prototypes(ctx, tax, cat_enc.counts, np.arange(len(seqs), dtype=np.int32),
                              sum_parents=False)
    q = H.encode(ctx, _dev_seqs(H, [r for _, r in reads]), counts=True, limbs=False, hi=False)
    print(f"descend tax_short n_cat={tax.n_cat} reads={len(reads)} beam=5", flush=True)
    res = H.descend(ctx, tree, q, beam=5, threshold=0.0, levels=6)
    torch.cuda.synchronize()
    got = format_lines(names, reads, res.path.cpu().numpy(), res.score.cpu().numpy(),
                       res.length.cpu().numpy(), res.exit_level.cpu().numpy())
    assert got == expected_lines()


def _synthetic(sizes, seq_len, n_reads, seed, dup_leaves=False, read_len=(150, 1600),
               with_truth=False):
    import hdc_amd as H
    tax, rd, roff, truth = H.make_taxonomy(sizes=sizes, seq_len=seq_len, n_reads=n_reads,
                                           read_len=read_len, seed=seed)
    leaf = np.frombuffer(tax.leaf_seqs, np.uint8).reshape(sizes[-1], seq_len).copy()
    if dup_leaves:  # every leaf equal to its first sibling: exact ties among siblings
        par = np.arange(sizes[-1]) * sizes[-2] // sizes[-1]  # the generator's parent rule
        first = {}
        for i in range(sizes[-1]):
            first.setdefault(par[i], i)
            leaf[i] = leaf[first[par[i]]]
    leaf_seqs = [leaf[i].tobytes().decode() for i in range(sizes[-1])]
    reads = [rd[roff[i]:roff[i + 1]].decode() for i in range(n_reads)]
    if with_truth:
        return tax, leaf_seqs, reads, truth
    return tax, leaf_seqs, reads


def _oracle_protos(D, k, tax, leaf_seqs):
    proto = np.zeros((tax.n_cat, D), np.int64)
    for c, s in zip(tax.ranks[-1], leaf_seqs):
        proto[c] = orc.compute(D, k, s)[0]
    for rank in reversed(tax.ranks[:-1]):
        for p in rank:
            kids = tax.child_idx[tax.child_off[p]:tax.child_off[p + 1]]
            if len(kids) == 1 and kids[0] == p:
                continue
            proto[p] = proto[kids].sum(axis=0)
    return proto.astype(np.int32)


@pytest.fixture(scope="module")
def synth():
    import hdc_amd as H
    import torch
    D, k = 10240, 12
    tax, leaf_seqs, reads, truth = _synthetic((4, 12, 24, 48, 96, 240), 1500, 160, seed=3,
                                              with_truth=True)
    # edge reads: empty, shorter than k, all N, lower case
    reads = reads + ["", "ACGTACGTAC", "N" * 300, reads[0].lower()]
    ctx = H.Context(D, k)
    leaf_enc = H.encode(ctx, _dev_seqs(H, leaf_seqs), counts=True, limbs=False, hi=False)
    tree = H.build_prototypes(ctx, tax, leaf_enc.counts, tax.ranks[-1])
    q = H.encode(ctx, _dev_seqs(H, reads), counts=True, limbs=False, hi=False)
    torch.cuda.synchronize()
    proto = _oracle_protos(D, k, tax, leaf_seqs)
    qc = q.counts.cpu().numpy()[:, :D]
    return ctx, tax, tree, q, proto, qc, truth


def test_prototypes_match_oracle(synth):
    ctx, tax, tree, q, proto, qc, truth = synth
    D = proto.shape[1]
    got = tree.proto.cpu().numpy()
    assert np.array_equal(got[:, :D], proto)
    assert not got[:, D:].any()
    lo = tree.lo.cpu().numpy().astype(np.int64)
    hi = tree.hi.cpu().numpy().astype(np.int64)
    assert np.array_equal(65536 * hi[:, :D] + lo[:, :D], proto.astype(np.int64))
    sq = (proto.astype(np.int64) ** 2).sum(axis=1)
    assert np.array_equal(tree.sqnorm.cpu().numpy(), sq)
    assert np.array_equal(tree.row_maxabs.cpu().numpy(), np.abs(proto.astype(np.int64)).max(axis=1))


def _run_both(synth, beam, thr, levels=6, q_maxabs=None):
    import hdc_amd as H
    import torch
    ctx, tax, tree, q, proto, qc, truth = synth
    print(f"descend synthetic beam={beam} thr={thr} levels={levels} q_maxabs={q_maxabs}",
          flush=True)
    res = H.descend(ctx, tree, q, beam=beam, threshold=thr, levels=levels, q_maxabs=q_maxabs)
    torch.cuda.synchronize()
    ref = orc.descend(proto, tax.child_off, tax.child_idx, tax.roots, qc, beam, thr, levels)
    got = (res.path.cpu().numpy(), res.score.cpu().numpy(), res.length.cpu().numpy(),
           res.exit_level.cpu().numpy())
    return got, ref


def _assert_same(got, ref):
    gp, gs, gl, ge = got
    rp, rs, rl, re_ = ref
    assert np.array_equal(gl, rl)
    assert np.array_equal(ge, re_)
    for r in range(len(gl)):
        n = gl[r]
        assert np.array_equal(gp[r, :n], rp[r, :n]), f"read {r}"
        assert np.array_equal(gs[r, :n].view(np.int32), rs[r, :n].view(np.int32)), f"read {r}"


@pytest.mark.parametrize("beam,thr", [(3, -1.0), (5, -1.0), (1, -1.0), (16, -1.0), (3, 0.0),
                                      (5, -0.6)])
def test_descend_matches_oracle(synth, beam, thr):
    """threshold -1 keeps every candidate above the cut (the integer-sum
    prototypes of high ranks score reads near -0.7 on identify's 2c-1 scale,
    below its 0.0 cut); 0.0 and -0.6 exercise early exits."""
    got, ref = _run_both(synth, beam, thr)
    _assert_same(got, ref)
    path, ex = got[0], got[3]
    truth = synth[6]
    n = len(truth)
    if thr == -1.0:
        assert (ex < 0).all()  # every read runs all six ranks (zero-norm reads score -1)
        if beam >= 3:
            assert (path[:n, 5] == truth).mean() > 0.9
    if thr == 0.0:
        assert ex[-4] == 0 and ex[-3] == 0 and ex[-2] == 0  # empty / short / all-N reads
        assert (ex == 0).sum() > n // 2


def test_descend_int8_and_int16_rows_match_oracle(synth):
    """By default, reads whose every count fits int8 are scored from int8 rows
    against two int8 digit planes of the prototypes, the others from their
    int16 rows; HVK_TUNE_DESCENT = 1 scores every read from int16 rows.  The
    synthetic reads (150-1,600 nt) hold both kinds, and both settings give
    the oracle's descent, with the int8 rows written by the dense level 0
    and by the conversion kernel of the vector level 0."""
    import hdc_amd as H
    ctx, tax, tree, q, proto, qc, truth = synth
    mx = np.abs(qc.astype(np.int32)).max(axis=1)
    assert (mx <= 127).any() and (mx > 128).any()
    try:
        for dense in (True, False):
            tree.use_dense = dense
            for mode in (1, 0):
                ctx.set_tuning(H.TUNE_DESCENT, mode)
                print(f"descend synthetic int8 rows mode={mode} dense_level0={dense}", flush=True)
                got, ref = _run_both(synth, 3, -1.0)
                _assert_same(got, ref)
    finally:
        ctx.set_tuning(H.TUNE_DESCENT, 0)
        tree.use_dense = True


def test_descend_int8_row_boundary_counts(synth):
    """Count rows given directly, with entries at the edges of int8 (-128 and
    127 fit, -129 and 128 do not) and far beyond it: a read takes the int8 rows
    exactly when every count fits, and every read's descent equals the
    oracle's, for both level-0 forms and both settings."""
    import hdc_amd as H
    import torch
    ctx, tax, tree, q, proto, qc, truth = synth
    D = proto.shape[1]
    rng = np.random.default_rng(5)
    base = qc[:8].astype(np.int32)
    groups = []
    for v in (-128, 127, 128, -129, 300, -20000):
        x = np.clip(base, -100, 100)
        x[:, rng.integers(0, D, 3)] = v
        groups.append(x)
    groups.append(np.clip(base, -127, 127))
    counts = np.concatenate(groups).astype(np.int16)
    n = len(counts)

    class Rows:
        pass

    rows = Rows()
    rows.n = n
    rows.counts = torch.zeros((n, ctx.ld), dtype=torch.int16, device="cuda")
    rows.counts[:, :D] = torch.from_numpy(counts).cuda()
    rows.sqnorm = torch.from_numpy((counts.astype(np.int64) ** 2).sum(axis=1)).cuda()
    rows.maxabs = torch.from_numpy(np.abs(counts.astype(np.int32)).max(axis=1)).cuda()
    fits = (counts.astype(np.int32) >= -128).all(axis=1) & (counts.astype(np.int32) <= 127).all(axis=1)
    assert fits[:16].all() and not fits[16:48].any() and fits[48:].all()
    ref = orc.descend(proto, tax.child_off, tax.child_idx, tax.roots, counts, 3, -1.0, 6)
    try:
        for dense in (True, False):
            tree.use_dense = dense
            for mode in (0, 1):
                ctx.set_tuning(H.TUNE_DESCENT, mode)
                print(f"descend boundary counts mode={mode} dense_level0={dense}", flush=True)
                res = H.descend(ctx, tree, rows, beam=3, threshold=-1.0, levels=6)
                torch.cuda.synchronize()
                _assert_same((res.path.cpu().numpy(), res.score.cpu().numpy(),
                              res.length.cpu().numpy(), res.exit_level.cpu().numpy()), ref)
    finally:
        ctx.set_tuning(H.TUNE_DESCENT, 0)
        tree.use_dense = True


def test_descend_wide_candidate_lists_match_oracle():
    """Candidate lists of 10 (wider than the 5 of a five-wave pass): the
    8-wide passes, two per list, give the oracle's paths and scores."""
    import hdc_amd as H
    import torch
    D, k = 10240, 12
    sizes = (2, 20, 40, 80, 160, 1600)  # spans 10, 2, 2, 2, 10
    tax, leaf_seqs, reads = _synthetic(sizes, 1500, 96, seed=11)
    assert np.diff(np.asarray(tax.child_off)).max() == 10
    ctx = H.Context(D, k)
    leaf_enc = H.encode(ctx, _dev_seqs(H, leaf_seqs), counts=True, limbs=False, hi=False)
    tree = H.build_prototypes(ctx, tax, leaf_enc.counts, tax.ranks[-1])
    q = H.encode(ctx, _dev_seqs(H, reads), counts=True, limbs=False, hi=False)
    torch.cuda.synchronize()
    proto = _oracle_protos(D, k, tax, leaf_seqs)
    qc = q.counts.cpu().numpy()[:, :D]
    for beam in (3, 5):
        print(f"descend wide lists beam={beam}", flush=True)
        res = H.descend(ctx, tree, q, beam=beam, threshold=-1.0, levels=6)
        torch.cuda.synchronize()
        ref = orc.descend(proto, tax.child_off, tax.child_idx, tax.roots, qc, beam, -1.0, 6)
        _assert_same((res.path.cpu().numpy(), res.score.cpu().numpy(),
                      res.length.cpu().numpy(), res.exit_level.cpu().numpy()), ref)
    ctx.close()


def test_descend_skips_zero_hi_limbs_exactly():
    """Rows with |p| >= 2^15 (a parent of 300 equal leaves) need the hi limb;
    leaf rows do not and skip it.  A candidate list mixing both, and the
    two tests of a zero hi row (per-row maxima, squared norms), give the
    oracle's paths and scores."""
    import hdc_amd as H
    import torch
    D, k = 10240, 12
    rng = np.random.default_rng(21)
    s1 = "".join(rng.choice(list("ACGT"), 1500))
    s2 = "".join(rng.choice(list("ACGT"), 1500))
    # 0: root; 1: A (300 leaves equal to s1); 2: B (one leaf s2); 3..302: A's
    # leaves; 303: B's leaf.  Leaves list themselves.
    n_leaf_a = 300
    n_cat = 3 + n_leaf_a + 1
    kids = [[1, 2], list(range(3, 3 + n_leaf_a)), [3 + n_leaf_a]] + \
        [[c] for c in range(3, n_cat)]
    off = np.concatenate([[0], np.cumsum([len(x) for x in kids])]).astype(np.int32)
    idx = np.concatenate(kids).astype(np.int32)
    ranks = [np.array([0]), np.array([1, 2]), np.arange(3, n_cat)]
    tax = H.Taxonomy(off, idx, [0], ranks)
    leaf_seqs = [s1] * n_leaf_a + [s2]
    reads = [s1[i:i + 400] for i in range(0, 1000, 100)] + \
        [s2[i:i + 700] for i in range(0, 800, 200)] + ["".join(rng.choice(list("ACGT"), 500))]
    ctx = H.Context(D, k)
    leaf_enc = H.encode(ctx, _dev_seqs(H, leaf_seqs), counts=True, limbs=False, hi=False)
    tree = H.build_prototypes(ctx, tax, leaf_enc.counts, ranks[-1])
    q = H.encode(ctx, _dev_seqs(H, reads), counts=True, limbs=False, hi=False)
    torch.cuda.synchronize()
    proto = _oracle_protos(D, k, tax, leaf_seqs)
    row_max = np.abs(proto.astype(np.int64)).max(axis=1)
    assert np.array_equal(tree.row_maxabs.cpu().numpy(), row_max)
    assert row_max[1] >= 32768 and row_max[2] < 32768 and row_max[3:].max() < 32768
    qc = q.counts.cpu().numpy()[:, :D]
    for use_rows in (True, False):
        tree.use_row_maxabs = use_rows
        for beam in (1, 3):
            print(f"descend hi-limb tree rows={use_rows} beam={beam}", flush=True)
            res = H.descend(ctx, tree, q, beam=beam, threshold=-1.0, levels=3)
            torch.cuda.synchronize()
            ref = orc.descend(proto, off, idx, [0], qc, beam, -1.0, 3)
            _assert_same((res.path.cpu().numpy(), res.score.cpu().numpy(),
                          res.length.cpu().numpy(), res.exit_level.cpu().numpy()), ref)
    ctx.close()


def test_descend_narrow_partial_sums(synth):
    """Claiming larger |q| (5000, 20000) forces the 8- and then the 2-element int32
    partials: same results."""
    base = _run_both(synth, 3, 0.0)
    for qm in (5000, 20000):
        got, ref = _run_both(synth, 3, 0.0, q_maxabs=qm)
        _assert_same(got, base[0])


def test_descend_ties_follow_std_sort():
    """Sibling leaves with equal sequences tie exactly; beam 5 over 5-leaf
    genera gives 25 candidates at the last rank, so the introsort branch of
    std::sort orders them."""
    import hdc_amd as H
    import torch
    D, k = 10240, 12
    tax, leaf_seqs, reads = _synthetic((3, 6, 12, 24, 48, 240), 1200, 120, seed=5,
                                       dup_leaves=True, read_len=(300, 1200))
    ctx = H.Context(D, k)
    leaf_enc = H.encode(ctx, _dev_seqs(H, leaf_seqs), counts=True, limbs=False, hi=False)
    tree = H.build_prototypes(ctx, tax, leaf_enc.counts, tax.ranks[-1])
    q = H.encode(ctx, _dev_seqs(H, reads), counts=True, limbs=False, hi=False)
    print(f"descend ties n_cat={tax.n_cat} reads={len(reads)} beam=5", flush=True)
    res = H.descend(ctx, tree, q, beam=5, threshold=-1.0, levels=6)
    torch.cuda.synchronize()
    proto = _oracle_protos(D, k, tax, leaf_seqs)
    qc = q.counts.cpu().numpy()[:, :D]
    ref = orc.descend(proto, tax.child_off, tax.child_idx, tax.roots, qc, 5, -1.0, 6)
    got = (res.path.cpu().numpy(), res.score.cpu().numpy(), res.length.cpu().numpy(),
           res.exit_level.cpu().numpy())
    _assert_same(got, ref)
    # the ties are real: with candidate lists reversed the oracle's result changes
    rev = np.concatenate([tax.child_idx[tax.child_off[c]:tax.child_off[c + 1]][::-1]
                          for c in range(tax.n_cat)])
    alt = orc.descend(proto, tax.child_off, rev, tax.roots, qc, 5, -1.0, 6)
    assert (got[3] < 0).all() and not np.array_equal(alt[0], ref[0])


def test_descend_rejects_bad_trees():
    """Host validation rejects malformed candidate lists and out-of-range limbs.
    Every device argument is a real buffer of the right shape, so a validation
    regression fails the status assertion instead of reading bad addresses."""
    import hdc_amd as H
    import ctypes
    import torch
    ctx = H.Context(10240, 12)
    n_cat, n_q, levels = 2, 4, 6
    lo = torch.zeros((n_cat, ctx.ld), dtype=torch.int16, device="cuda")
    hi = torch.zeros_like(lo)
    sq = torch.ones(n_cat, dtype=torch.int64, device="cuda")
    qc = torch.zeros((n_q, ctx.ld), dtype=torch.int16, device="cuda")
    qn = torch.ones(n_q, dtype=torch.int64, device="cuda")
    path = torch.zeros((n_q, levels), dtype=torch.int32, device="cuda")
    score = torch.zeros((n_q, levels), dtype=torch.float32, device="cuda")
    length = torch.zeros(n_q, dtype=torch.int32, device="cuda")
    exit_level = torch.zeros(n_q, dtype=torch.int32, device="cuda")
    t = H.Tree()
    t.n_cat = n_cat
    t.ld = ctx.ld
    off = np.array([0, 1, 5], np.int32)
    idx = np.array([1, 1, 7, 0, 0], np.int32)  # 7 is out of range
    roots = np.array([0], np.int32)
    t.child_off, t.child_idx, t.roots, t.n_roots = off.ctypes.data, idx.ctypes.data, \
        roots.ctypes.data, 1
    t.d_lo, t.d_hi, t.d_sqnorm = lo.data_ptr(), hi.data_ptr(), sq.data_ptr()
    t.max_abs_limb = 100
    args = [H._ptr(qc), H._ptr(qn), n_q, ctx.ld, 100, 3, 0.0, levels, H._ptr(path), H._ptr(score),
            H._ptr(length), H._ptr(exit_level), H._stream()]
    rc = H.lib().hvk_descend(ctx.h, ctypes.byref(t), *args)
    assert rc == H.HVK_ERR_ARG
    idx[2] = 1
    t.max_abs_limb = 40000
    rc = H.lib().hvk_descend(ctx.h, ctypes.byref(t), *args)
    assert rc == H.HVK_ERR_RANGE
    torch.cuda.synchronize()
    ctx.close()


def test_dense_level0_equals_vector_path(synth):
    """Level 0 on the matrix cores (int8 digit GEMM) and on the vector path give
    identical descents, with the prototypes split into 2, 3 and 4 digits (a
    larger claimed max |p| than the true one only adds digits)."""
    import hdc_amd as H
    import torch
    ctx, tax, tree, q, proto, qc, truth = synth
    true_max = tree.max_abs_proto
    assert 0 < true_max <= 32639  # two digits
    outs = []
    for dense, claim in ((False, true_max), (True, true_max), (True, 40000), (True, 9_000_000)):
        tree.use_dense = dense
        tree.max_abs_proto = claim
        print(f"descend synthetic dense_level0={dense} max_abs_proto={claim}", flush=True)
        r = H.descend(ctx, tree, q, beam=3, threshold=-1.0, levels=6)
        torch.cuda.synchronize()
        outs.append((r.path.cpu().numpy(), r.score.cpu().numpy().view(np.int32),
                     r.length.cpu().numpy(), r.exit_level.cpu().numpy()))
    tree.use_dense = True
    tree.max_abs_proto = true_max
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a, b)
