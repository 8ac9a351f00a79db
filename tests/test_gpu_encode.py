"""GPU encoder parity: hvk_encode vs the C oracle (HVCache::compute)."""
import numpy as np
import pytest

import oracle_py as orc
import seqgen

pytestmark = pytest.mark.gpu

# One wave per sequence (k in {10, 12}, 129..320 words of 32 dims): (8192, 10)
# four words per lane, (10240, 12) and (10000, 12) five (the fifth word of
# lanes 57..63 past D=10000), (5000, 12) four words with lanes past the row.
# The same (D, k) with legacy=True and (10000, 5), (20000, 7) run the
# workgroup-per-sequence pair-table kernel ((10000, 5) three parts, (8192, 10)
# four, (10240, 12) three, (20000, 7) one).  (20000, 12): the pair table does
# not fit in LDS, so the single-table kernel runs.
CASES = [(10000, 5, False), (8192, 10, False), (8192, 10, True), (10240, 12, False),
         (10240, 12, True), (10000, 12, False), (5000, 12, False), (20000, 7, False),
         (20000, 12, False)]


def _encode(dim, k, seqs, legacy=False, **kw):
    import hdc_amd as H
    import torch
    ctx = H.Context(dim, k)
    if legacy:
        ctx.set_tuning(H.TUNE_ENCODER, 1)
    buf, off = H.concat(seqs)
    dev = H.DeviceSeqs(buf, off)
    enc = H.encode(ctx, dev, **kw)
    torch.cuda.synchronize()
    return ctx, enc


@pytest.mark.parametrize("dim,k,legacy", CASES)
def test_counts_bit_exact(dim, k, legacy):
    seqs = seqgen.mixed_set(1000 + k, 24, 1, 400)
    seqs += ["", "A" * (k - 1), "ACGT" * 3, "N" * 50, "acgtacgtacgtacgtacgt", "ACGTN" * 40]
    # chunk edges of both kernels (512 and 1024 k-mer positions)
    rng = np.random.default_rng(dim + k)
    seqs += [seqgen.random_seq(rng, n + k - 1) for n in (511, 512, 513, 1024, 1025, 1600)]
    print(f"encode D={dim} k={k} legacy={legacy} n={len(seqs)}", flush=True)
    ctx, enc = _encode(dim, k, seqs, legacy)
    counts = enc.counts.cpu().numpy()
    lo = enc.lo.cpu().numpy().astype(np.int32)
    hi = enc.hi.cpu().numpy().astype(np.int32)
    sq = enc.sqnorm.cpu().numpy()
    mx = enc.maxabs.cpu().numpy()
    nk = enc.nkmers.cpu().numpy()
    for i, s in enumerate(seqs):
        ref, n = orc.compute(dim, k, s)
        assert np.array_equal(counts[i, :dim], ref), f"seq {i} len {len(s)}"
        assert not counts[i, dim:].any()
        assert np.array_equal(256 * hi[i, :dim] + lo[i, :dim], ref.astype(np.int32))
        assert sq[i] == int((ref.astype(np.int64) ** 2).sum())
        assert mx[i] == int(np.abs(ref.astype(np.int32)).max())
        assert nk[i] == n


@pytest.mark.parametrize("legacy", [False, True])
def test_long_sequences_planes(legacy):
    """> 2048 k-mers (16 counter planes) and the int32 oracle sum."""
    rng = np.random.default_rng(7)
    seqs = [seqgen.random_seq(rng, 5000), "A" * 3000, seqgen.random_seq(rng, 2100)]
    ctx, enc = _encode(10240, 12, seqs, legacy)
    counts = enc.counts.cpu().numpy()
    for i, s in enumerate(seqs):
        ref32, _ = orc.compute32(10240, 12, s)
        assert np.array_equal(counts[i, :10240].astype(np.int32), ref32)


def test_generated_references_exact():
    """Sample of the synthetic workload's references (1,300-1,600 nt)."""
    import hdc_amd as H
    spec = H.make_spec(n_ancestors=4, desc_per_ancestor=4, anc_len_min=1300, anc_len_max=1600,
                       div_min=0.02, div_max=0.08, n_reads=16, read_len_min=150, read_len_max=300,
                       err_min=0.005, err_max=0.02)
    w = H.generate(spec)
    refs = [w.refs[w.ref_offsets[i]:w.ref_offsets[i + 1]].decode() for i in range(16)]
    reads = [w.reads[w.read_offsets[i]:w.read_offsets[i + 1]].decode() for i in range(16)]
    ctx, enc = _encode(10240, 12, refs + reads)
    counts = enc.counts.cpu().numpy()
    for i, s in enumerate(refs + reads):
        ref, _ = orc.compute(10240, 12, s)
        assert np.array_equal(counts[i, :10240], ref), i


def test_base_hv_matches_oracle():
    import hdc_amd as H
    for dim in (10000, 10240):
        ctx = H.Context(dim, 5)
        for ch in "ACGTN":
            assert np.array_equal(ctx.base_hv(ch), orc.base_hv(dim, ch))


@pytest.mark.parametrize("dim,k,legacy", [(10240, 12, False), (10240, 12, True),
                                          (10000, 5, False)])
def test_clamp_and_residual_outputs(dim, k, legacy):
    """clamp8 = clip(count, -127, 127); resid1 = sum |count - clamp8|."""
    rng = np.random.default_rng(21)
    seqs = [seqgen.random_seq(rng, 1500), "A" * 400, seqgen.random_seq(rng, 200), "ACGTAC" * 100]
    ctx, enc = _encode(dim, k, seqs, legacy, clamp=True)
    c = enc.counts.cpu().numpy().astype(np.int32)[:, :dim]
    cl = enc.clamp8.cpu().numpy().astype(np.int32)
    want = np.clip(c, -127, 127)
    assert np.array_equal(cl[:, :dim], want)
    assert not cl[:, dim:].any()
    assert np.array_equal(enc.resid1.cpu().numpy(), np.abs(c - want).sum(1))
    assert enc.resid1.cpu().numpy()[1] > 0  # poly-A saturates


@pytest.mark.parametrize("dim,k,legacy", [(10000, 5, False), (10240, 12, False),
                                          (10240, 12, True)])
def test_windows_match_substring_compute(dim, k, legacy):
    """hvk_encode_windows == compute() of each window's substring."""
    import hdc_amd as H
    import torch
    seqs = seqgen.mixed_set(55, 5, 300, 700)
    buf, off = H.concat(seqs)
    ctx = H.Context(dim, k)
    if legacy:
        ctx.set_tuning(H.TUNE_ENCODER, 1)
    dev = H.DeviceSeqs(buf, off)
    starts = []
    for i, s in enumerate(seqs):
        starts += [int(off[i]) + x for x in range(0, len(s) - 125 + 1, 10)]
    starts = np.array(starts[:200], np.int64)
    enc = H.encode_windows(ctx, dev, starts, 125)
    torch.cuda.synchronize()
    got = enc.counts.cpu().numpy()[:, :dim]
    text = buf.decode()
    for j in range(0, len(starts), 7):
        ref, _ = orc.compute(dim, k, text[starts[j]:starts[j] + 125])
        assert np.array_equal(got[j], ref), j
