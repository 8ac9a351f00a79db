"""Prototype database file on the device: an encoded reference set saved with
hvk_db_write and loaded with hvk_db_upload is identical, and searching it
gives the same top-k as the freshly encoded set."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_save_load_search(tmp_path):
    import hdc_amd as H
    import torch
    spec = H.make_spec(n_ancestors=6, desc_per_ancestor=20, anc_len_min=1300, anc_len_max=1600,
                       div_min=0.02, div_max=0.08, n_reads=300, read_len_min=150, read_len_max=300,
                       err_min=0.005, err_max=0.02)
    w = H.generate(spec)
    ctx = H.Context(10240, 12)
    db = H.encode(ctx, H.DeviceSeqs(w.refs, w.ref_offsets), counts=True, limbs=True, hi=True,
                  clamp=True)
    q = H.encode(ctx, H.DeviceSeqs(w.reads, w.read_offsets), counts=False, limbs=True, hi=True,
                 clamp=True)
    torch.cuda.synchronize()
    path = str(tmp_path / "refs.hvk")
    names = b"\n".join(b"ref%d" % i for i in range(db.n))
    H.db_save(path, ctx, db, names)
    print("dbfile save/load", path, db.n, flush=True)
    f = H.DbFile(path)
    assert f.n == db.n and f.names() == names
    db2 = f.load(ctx)
    torch.cuda.synchronize()
    for key in ("counts", "lo", "hi", "clamp8", "sqnorm", "maxabs", "resid1", "nkmers"):
        assert torch.equal(getattr(db, key), getattr(db2, key)), key
    a = H.search_mfma_bounded(ctx, q, db, 5)
    b = H.search_mfma_bounded(ctx, q, db2, 5)
    torch.cuda.synchronize()
    assert torch.equal(a.idx, b.idx) and torch.equal(a.score, b.score)
    f.close()
    other = H.Context(8192, 10)
    with pytest.raises(H.HvkError):
        H.DbFile(path).load(other)
