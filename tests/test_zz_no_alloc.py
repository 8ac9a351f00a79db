"""hvk_search_mfma_bounded_noalloc (the allocation-free bounded search, its
buffers in a fixed device array) gives exactly hvk_search_mfma_bounded's
results, called through the C ABI twice in a row, the second call smaller
(the array keeps the first call's contents), and rejects sizes past its
array without touching the device."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIM, K = 10240, 12


def _run(fn, ctx, q, r, k):
    import hdc_amd as H
    o = H._topk_bufs(q.n, k)
    bq, br = H.bundles(q), H.bundles(r)
    rc = fn(ctx.h, ctypes.byref(bq), ctypes.byref(br), ctx.ld, k, H._ptr(o.idx), H._ptr(o.score),
            H._ptr(o.cos64), H._ptr(o.n_rescored), H._stream())
    return rc, o


def _slice(e, n):
    import hdc_amd as H
    t = lambda x: None if x is None else x[:n]
    return H.Encoded(n, t(e.counts), t(e.lo), t(e.hi), t(e.sqnorm), t(e.maxabs), t(e.nkmers),
                     t(e.clamp8), t(e.resid1))


def test_noalloc_equals_bounded_twice_second_smaller():
    import hdc_amd as H
    import oracle_py as orc
    import torch
    spec = H.make_spec(seed=9, n_ancestors=12, desc_per_ancestor=40, anc_len_min=1300,
                       anc_len_max=1600, div_min=0.002, div_max=0.03, n_reads=1200,
                       read_len_min=150, read_len_max=300, err_min=0.005, err_max=0.02)
    w = H.generate(spec)
    ctx = H.Context(DIM, K)
    db = H.encode(ctx, H.DeviceSeqs(w.refs, w.ref_offsets), clamp=True)
    reads = H.encode(ctx, H.DeviceSeqs(w.reads, w.read_offsets), clamp=True)
    torch.cuda.synchronize()
    L = H.lib()
    for n, k in ((reads.n, 10), (300, 5)):  # second call smaller
        print(f"noalloc call n={n} k={k} nr={db.n}", flush=True)
        q = _slice(reads, n)
        rc0, want = _run(L.hvk_search_mfma_bounded, ctx, q, db, k)
        rc1, got = _run(L.hvk_search_mfma_bounded_noalloc, ctx, q, db, k)
        assert rc0 == H.HVK_OK and rc1 == H.HVK_OK, (rc0, rc1, L.hvk_last_error())
        torch.cuda.synchronize()
        assert torch.equal(got.idx, want.idx)
        assert torch.equal(got.score, want.score)
        assert torch.equal(got.cos64, want.cos64)
        assert torch.equal(got.n_rescored, want.n_rescored)
    # and the smaller call against the oracle's FP64 brute force
    qc = reads.counts.cpu().numpy()[:300, :DIM]
    rcnt = db.counts.cpu().numpy()[:, :DIM]
    ref_idx, _ = orc.topk16(qc, rcnt, 5)
    assert np.array_equal(got.idx.cpu().numpy(), ref_idx)
    ctx.close()


def test_noalloc_rejects_sizes_past_its_array():
    """160,000 queries x 4,096 references need ~82 MB of candidate lists,
    past the 64 MiB array: HVK_ERR_RANGE before any launch.  Every operand is
    a real (zero) device buffer of its full size."""
    import hdc_amd as H
    import torch
    ctx = H.Context(DIM, K)
    L = H.lib()
    q = H._alloc(160_000, ctx.ld, counts=False, limbs=True, hi=False, clamp=True)
    r = H._alloc(4096, ctx.ld, counts=False, limbs=True, hi=False, clamp=True)
    for e in (q, r):
        for t in (e.lo, e.clamp8, e.sqnorm, e.maxabs, e.resid1):
            t.zero_()
    bq, br = H.bundles(q, False), H.bundles(r, False)
    o = H._topk_bufs(q.n, 5)
    print("noalloc rejection nq=160000 nr=4096", flush=True)
    rc = L.hvk_search_mfma_bounded_noalloc(ctx.h, ctypes.byref(bq), ctypes.byref(br), ctx.ld, 5,
                                           H._ptr(o.idx), H._ptr(o.score), None, None, H._stream())
    assert rc == H.HVK_ERR_RANGE
    assert b"scratch" in L.hvk_last_error()
    rc = L.hvk_search_mfma_bounded_noalloc(ctx.h, ctypes.byref(bq), ctypes.byref(br), ctx.ld, 30,
                                           H._ptr(o.idx), H._ptr(o.score), None, None, H._stream())
    assert rc == H.HVK_ERR_ARG  # k past the candidate lists
    torch.cuda.synchronize()
    del q, r, o
    ctx.close()
