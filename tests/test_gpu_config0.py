"""The CPU-parity workload on the GPU: D=8192, k=10; 20 random ancestors of
1,500 nt, each mutated into 50 descendants at 3% substitution and 1% indels
(1,000 references); 500 reads of 250 nt sampled from the references with 1%
substitutions; top-1 cosine argmax (and top-10), as the reference's own
GPU-vs-CPU check does (tests/cosine_comparison_test.cpp:121-159, top-1 and
top-10 agreement).  Every search form must give the FP64 oracle's indices,
and scores within 1e-6 relative; the encoding is bit-exact."""
import numpy as np
import pytest

import oracle_py as orc

pytestmark = pytest.mark.gpu

DIM, K = 8192, 10
COS_REL_TOL = 1e-6


@pytest.fixture(scope="module")
def config0():
    import hdc_amd as H
    import torch
    spec = H.make_spec()  # the generator's defaults are this workload
    w = H.generate(spec)
    assert len(w.ref_offsets) - 1 == 1000 and len(w.read_offsets) - 1 == 500
    ctx = H.Context(DIM, K)
    refs = H.encode(ctx, H.DeviceSeqs(w.refs, w.ref_offsets), clamp=True)
    reads = H.encode(ctx, H.DeviceSeqs(w.reads, w.read_offsets), clamp=True)
    torch.cuda.synchronize()
    return ctx, w, refs, reads


def test_config0_encoding_bit_exact(config0):
    ctx, w, refs, reads = config0
    qc = reads.counts.cpu().numpy()[:, :DIM]
    rc = refs.counts.cpu().numpy()[:, :DIM]
    for i in (0, 123, 499):
        s = w.reads[w.read_offsets[i]:w.read_offsets[i + 1]].decode()
        want, _ = orc.compute(DIM, K, s)
        assert np.array_equal(qc[i], want)
    for i in (0, 500, 999):
        s = w.refs[w.ref_offsets[i]:w.ref_offsets[i + 1]].decode()
        want, _ = orc.compute(DIM, K, s)
        assert np.array_equal(rc[i], want)


@pytest.mark.parametrize("form", ["plain", "mfma", "bounded"])
@pytest.mark.parametrize("k", [1, 10])
def test_config0_topk_matches_oracle(config0, form, k):
    import hdc_amd as H
    import torch
    ctx, w, refs, reads = config0
    print(f"config0 form={form} k={k}", flush=True)
    qc = reads.counts.cpu().numpy()[:, :DIM]
    rc = refs.counts.cpu().numpy()[:, :DIM]
    ref_idx, ref_cos = orc.topk16(qc, rc, k)
    if form == "plain":
        res = H.search_plain(ctx, reads, refs, k)
    elif form == "mfma":
        res = H.search_mfma(ctx, reads, refs, k)  # two limbs where |count| > 127
    else:
        res = H.search_mfma_bounded(ctx, reads, refs, k)
    torch.cuda.synchronize()
    assert np.array_equal(res.idx.cpu().numpy(), ref_idx)
    c64 = res.cos64.cpu().numpy()
    assert np.allclose(c64, ref_cos, rtol=1e-12, atol=1e-15)
    sc = res.score.cpu().numpy().astype(np.float64)
    assert (np.abs(sc - ref_cos) / np.abs(ref_cos)).max() <= COS_REL_TOL
    if k == 1:  # reads find their source reference (or an exact tie of it)
        hit = (res.idx.cpu().numpy()[:, 0] == w.truth).mean()
        assert hit > 0.9
