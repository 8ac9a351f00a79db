#!/usr/bin/env python3
"""Writes tests/golden/ref_prims_d10000_k5.npz: the reference's own host
primitives (src/HV.cpp:9-216, src/HV.h:21-106, src/encoder.cpp:47-74) on
fixed random inputs, run through oracle/_ref/libref.so (built by
`make -C oracle ref` from the reference checkout at its configuration
HV_DIM 10000, KMER_SIZE 5).  The drop-in's primitives (host/HV.cpp,
host/encoder.cpp) are compared with these vectors by tests/test_cpp_prims.py.

Contents: empty_hv; get_base_hv of A C G T N x; mult for 12 shifts (0, word
and ring boundaries, past D); superpose_batch from a nonzero accumulator for
batches of 1, 7 and 16; replace_batch for batches of 1 and 10; the size of
HVCache::kmers, the SHA-256 of all 1,024 k-mer HVs in lexicographic order
and 24 of them in full.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle_py as orc  # noqa: E402

SHIFTS = [0, 1, 5, 63, 64, 65, 127, 5000, 9935, 9999, 10000, 10047]


def main():
    R = orc.reference()
    if R is None:
        sys.exit("oracle/_ref/libref.so missing: run make -C oracle ref")
    D, k, _, _ = orc.ref_config()
    W = (D + 63) // 64
    tail = (1 << (D - 64 * (W - 1))) - 1
    rng = np.random.default_rng(77)

    def rand_hv(n):
        h = rng.integers(0, 2**63, (n, W), dtype=np.uint64) * np.uint64(2) + \
            rng.integers(0, 2, (n, W), dtype=np.uint64)
        h[:, -1] &= np.uint64(tail)
        return np.ascontiguousarray(h)

    empty = np.zeros(W, np.uint64)
    R.ref_empty_hv(empty.ctypes.data)
    syms = "ACGTNx"
    base = np.zeros((len(syms), W), np.uint64)
    for i, c in enumerate(syms):
        R.ref_base_hv(c.encode(), base[i].ctypes.data)
    ml, mr = rand_hv(len(SHIFTS)), rand_hv(len(SHIFTS))
    mout = np.zeros_like(ml)
    for i, s in enumerate(SHIFTS):
        R.ref_mult(ml[i].ctypes.data, mr[i].ctypes.data, mout[i].ctypes.data, s)
    sizes = [1, 7, 16]
    sp_hvs = rand_hv(sum(sizes))
    sp_acc0 = rng.integers(-300, 300, (len(sizes), D)).astype(np.int16)
    sp_out = sp_acc0.copy()
    o = 0
    for i, n in enumerate(sizes):
        row = np.ascontiguousarray(sp_out[i])
        R.ref_superpose_batch(row.ctypes.data, np.ascontiguousarray(sp_hvs[o:o + n]).ctypes.data, n)
        sp_out[i] = row
        o += n
    rsizes = [1, 10]
    rp_old, rp_new = rand_hv(sum(rsizes)), rand_hv(sum(rsizes))
    rp_acc0 = rng.integers(-300, 300, (len(rsizes), D)).astype(np.int16)
    rp_out = rp_acc0.copy()
    o = 0
    for i, n in enumerate(rsizes):
        row = np.ascontiguousarray(rp_out[i])
        R.ref_replace_batch(row.ctypes.data, np.ascontiguousarray(rp_old[o:o + n]).ctypes.data,
                            np.ascontiguousarray(rp_new[o:o + n]).ctypes.data, n)
        rp_out[i] = row
        o += n
    allk = np.zeros((4 ** k, W), np.uint64)
    n_entries = R.ref_kmers_all(allk.ctypes.data)
    pick = np.array(sorted(set([0, 1, 2, 3, 4 ** k - 1] + list(rng.integers(0, 4 ** k, 19)))))
    np.savez_compressed(
        os.path.join(HERE, "ref_prims_d10000_k5.npz"),
        config=np.array([D, k], np.int64), empty=empty, base_syms=np.array(list(syms)).astype("U"),
        base=base, mult_shift=np.array(SHIFTS, np.int64), mult_l=ml, mult_r=mr, mult_out=mout,
        sp_sizes=np.array(sizes, np.int64), sp_hvs=sp_hvs, sp_acc0=sp_acc0, sp_out=sp_out,
        rp_sizes=np.array(rsizes, np.int64), rp_old=rp_old, rp_new=rp_new, rp_acc0=rp_acc0,
        rp_out=rp_out, kmers_size=np.array(n_entries, np.int64),
        kmers_sha256=np.array(hashlib.sha256(allk.tobytes()).hexdigest()).astype("U"),
        kmers_pick=pick.astype(np.int64), kmers_pick_hv=allk[pick])
    print("wrote", os.path.join(HERE, "ref_prims_d10000_k5.npz"), "kmers", n_entries)


if __name__ == "__main__":
    main()
