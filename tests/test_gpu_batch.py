"""Batching form (hvk_batch_*): length-ordered chunks replayed from hipGraphs
must give exactly the results of the direct path (hvk_encode_ex +
hvk_search_mfma_bounded), and those equal the oracle's FP64 brute force.

Reads of uneven length (log-uniform 100-1,600 nt, clamped to the source
reference) plus edge reads (empty, shorter than k, all N) go through several
chunk sizes, both search forms, first call (direct run + capture) and later
calls (graph replay, replayed several times with eager work and
synchronisations between), and the autotuner.
"""
import numpy as np
import pytest

import oracle_py as orc

pytestmark = pytest.mark.gpu

DIM, K, TOPK = 10240, 12, 5
_HOST_READS = []  # the fixture's reads, for tests that upload them again


@pytest.fixture(scope="module")
def work():
    import hdc_amd as H
    import torch
    spec = H.make_spec(n_ancestors=10, desc_per_ancestor=20, anc_len_min=1300, anc_len_max=1600,
                       div_min=0.02, div_max=0.08, n_reads=1500, read_len_min=100,
                       read_len_max=1600, read_len_log_uniform=1, err_min=0.01, err_max=0.01)
    w = H.generate(spec)
    reads = [w.reads[w.read_offsets[i]:w.read_offsets[i + 1]] for i in range(len(w.read_offsets) - 1)]
    reads += [b"", b"ACGTACGTAC", b"N" * 400, reads[3].lower()]
    buf, off = H.concat([r.decode() for r in reads])
    _HOST_READS[:] = reads
    ctx = H.Context(DIM, K)
    db = H.encode(ctx, H.DeviceSeqs(w.refs, w.ref_offsets), counts=True, limbs=True, hi=True,
                  clamp=True)
    seqs = H.DeviceSeqs(buf, off)
    q = H.encode(ctx, seqs, counts=True, limbs=True, hi=True, clamp=True)
    direct = H.search_mfma_bounded(ctx, q, db, TOPK)
    torch.cuda.synchronize()
    return ctx, db, seqs, q, direct


def _same(a, b):
    import torch
    torch.cuda.synchronize()
    assert torch.equal(a.idx, b.idx)
    assert torch.equal(a.score, b.score)
    assert torch.equal(a.cos64, b.cos64)


def test_direct_path_matches_oracle(work):
    ctx, db, seqs, q, direct = work
    qc = q.counts.cpu().numpy()[:, :DIM]
    rc = db.counts.cpu().numpy()[:, :DIM]
    ref_idx, ref_cos = orc.topk16(qc, rc, TOPK)
    assert np.array_equal(direct.idx.cpu().numpy(), ref_idx)


@pytest.mark.parametrize("chunk,form,splits,shape", [(256, 0, 0, 0), (512, 0, 2, 0), (1000, 0, 8, 0),
                                                     (4096, 0, 0, 0), (700, 1, 4, 0), (512, 0, 0, 1),
                                                     (768, 0, 2, 2), (512, 0, 0, 3), (1024, 0, 4, 4),
                                                     (600, 0, 0, 5)])
def test_batch_equals_direct(work, chunk, form, splits, shape):
    import hdc_amd as H
    import torch
    ctx, db, seqs, q, direct = work
    print(f"batch chunk={chunk} form={form} splits={splits} shape={shape} n={seqs.n}", flush=True)
    b = H.Batch(DIM, K, chunk)
    b.config(form, splits, chunk, shape)
    dbb = H.bundles(db)
    out = b.classify(seqs, dbb, TOPK)  # direct runs + captures
    _same(out, direct)
    n_chunks = (seqs.n + chunk - 1) // chunk
    assert b.get_config()["graphs"] == n_chunks
    for rep in range(3):  # replays, with eager work and syncs in between
        out.idx.fill_(-7)
        out.score.fill_(0.0)
        torch.cuda.synchronize()
        _ = (q.sqnorm * 2).sum().item()
        out = b.classify(seqs, dbb, TOPK, out=out)
        _same(out, direct)
    assert b.get_config()["graphs"] == n_chunks  # nothing re-captured
    b.close()


def test_batch_new_outputs_and_k(work):
    """Fresh output buffers replay the same graphs (the graphs touch only
    batch-owned buffers); another k is other graphs; results stay exact."""
    import hdc_amd as H
    ctx, db, seqs, q, direct = work
    b = H.Batch(DIM, K, 512)
    dbb = H.bundles(db)
    o1 = b.classify(seqs, dbb, TOPK)
    n_graphs = b.get_config()["graphs"]
    o2 = b.classify(seqs, dbb, TOPK)  # fresh output buffers: same graphs
    _same(o1, direct)
    _same(o2, direct)
    assert b.get_config()["graphs"] == n_graphs
    d10 = H.search_mfma_bounded(ctx, q, db, 10)
    o3 = b.classify(seqs, dbb, 10)
    _same(o3, d10)
    d40 = H.search_mfma_bounded(ctx, q, db, 40)  # past the candidate lists
    o4 = b.classify(seqs, dbb, 40)
    _same(o4, d40)
    o1 = b.classify(seqs, dbb, TOPK, out=o1)  # replay of the first graphs
    _same(o1, direct)
    b.close()


def test_batch_streaming_fresh_buffers_bounded_graphs(work):
    """A stream of calls, each with freshly uploaded reads (a copy of the same
    sequences, then a permutation of them) and fresh outputs: the graph count
    stays at the number of chunks and every result equals the direct path."""
    import hdc_amd as H
    import torch
    ctx, db, seqs, q, direct = work
    b = H.Batch(DIM, K, 512)
    dbb = H.bundles(db)
    n_chunks = (seqs.n + 511) // 512
    raw = list(_HOST_READS)
    for rep in range(3):
        order = np.arange(len(raw)) if rep == 0 else np.random.default_rng(rep).permutation(len(raw))
        buf, off = H.concat([raw[i].decode() for i in order])
        fresh = H.DeviceSeqs(buf, off)
        out = b.classify(fresh, dbb, TOPK)
        torch.cuda.synchronize()
        inv = torch.from_numpy(np.argsort(order)).cuda()
        assert torch.equal(out.idx[inv], direct.idx)
        assert torch.equal(out.cos64[inv], direct.cos64)
        assert b.get_config()["graphs"] <= 2 * n_chunks
    b.close()


def test_batch_autotune(work):
    import hdc_amd as H
    ctx, db, seqs, q, direct = work
    b = H.Batch(DIM, K, 512)
    dbb = H.bundles(db)
    out = H._topk_bufs(seqs.n, TOPK)
    table = []
    ms = b.autotune(seqs, dbb, TOPK, out, sample=1024, table=table)
    cfg = b.get_config()
    print("autotune", cfg, ms, flush=True)
    assert ms > 0 and cfg["chunk"] >= 256
    assert {r["shape"] for r in table} == set(range(6))  # every tile shape was timed
    assert min(r["ms_per_mread"] for r in table) == pytest.approx(ms)
    out = b.classify(seqs, dbb, TOPK, out=out)
    _same(out, direct)
    b.close()


def test_batch_descend_equals_direct():
    """hvk_batch_descend (length-ordered chunks of 256, each encoded and
    descended from a captured graph) gives hvk_encode_ex + hvk_descend's paths,
    scores (bit for bit), lengths and exit levels, on the first call and on a
    replay into fresh outputs; reads include empty, short and all-N ones."""
    import hdc_amd as H
    import torch
    D, k = 10240, 12
    tax, rd, roff, truth = H.make_taxonomy(sizes=(4, 12, 24, 48, 96, 240), seq_len=1500,
                                           n_reads=700, read_len=(150, 1600), seed=4)
    reads = [rd[roff[i]:roff[i + 1]].decode() for i in range(len(roff) - 1)]
    reads = reads[:300] + ["", "ACGTACGTAC", "N" * 300] + reads[300:]
    buf, off = H.concat(reads)
    seqs = H.DeviceSeqs(buf, off)
    ctx = H.Context(D, k)
    leaf_enc = H.encode(ctx, H.DeviceSeqs(tax.leaf_seqs, tax.leaf_offsets), counts=True,
                        limbs=False, hi=False)
    tree = H.build_prototypes(ctx, tax, leaf_enc.counts, tax.ranks[-1])
    q = H.encode(ctx, seqs, counts=True, limbs=False, hi=False)
    b = H.Batch(D, k, chunk=256)
    for beam, thr in ((3, -1.0), (5, 0.0)):
        print(f"batch descend beam={beam} thr={thr} reads={len(reads)}", flush=True)
        want = H.descend(ctx, tree, q, beam=beam, threshold=thr, levels=6)
        for rep in range(2):
            got = b.descend(tree, seqs, beam=beam, threshold=thr, levels=6)
            torch.cuda.synchronize()
            assert torch.equal(got.length, want.length)
            assert torch.equal(got.exit_level, want.exit_level)
            for r in range(len(reads)):
                n = int(want.length[r])
                assert torch.equal(got.path[r, :n], want.path[r, :n]), f"read {r}"
                assert torch.equal(got.score[r, :n].view(torch.int32),
                                   want.score[r, :n].view(torch.int32)), f"read {r}"
    b.close()
    ctx.close()
