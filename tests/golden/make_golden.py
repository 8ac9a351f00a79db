#!/usr/bin/env python3
"""Writes tests/golden/ref_d10000_k5.npz from the reference's own CPU sources.

Runs oracle/_ref/libref.so (built by `make -C oracle ref` from the reference
checkout: src/HV.cpp + src/encoder.cpp at their compile-time configuration
HV_DIM 10000, KMER_SIZE 5, window 125, step 10, NORM_COSINE) and stores its
outputs for fixed inputs, so the oracle stays pinned where the reference
checkout is absent.  Inputs: sequences of EXAMPLE_READS.fasta (the
reference's data file) when present, low-complexity 30-mers, random
sequences with IUPAC/N/lower-case noise.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle_py as orc  # noqa: E402
import seqgen  # noqa: E402

FASTA = "/root/reference/EXAMPLE_READS.fasta"


def read_fasta(path, n):
    out, cur = [], []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith(">"):
                if cur:
                    out.append("".join(cur))
                    cur = []
                    if len(out) == n:
                        break
            else:
                cur.append(line)
    if cur and len(out) < n:
        out.append("".join(cur))
    return out


def main():
    import ctypes
    R = orc.reference()
    if R is None:
        sys.exit("oracle/_ref/libref.so missing: run make -C oracle ref")
    D, k, window, step = orc.ref_config()
    W = (D + 63) // 64
    rng = np.random.default_rng(2024)
    # sequences with N / IUPAC codes from the reference's own example reads
    fasta = read_fasta(FASTA, 1000) if os.path.exists(FASTA) else []
    noisy = [s for s in fasta if any(c not in "ACGT" for c in s)][:2]
    clean = [s for s in fasta if all(c in "ACGT" for c in s)][:1]
    low = ["A" * 30, "AAG" * 10, "GAGAA" * 6, "GA" * 15, "GAGGA" * 6, "GGA" * 10, "GGAGG" * 6,
           "G" * 30]
    rand = seqgen.mixed_set(77, 6, 1, 300)
    seqs = noisy + clean + low + rand + ["", "ACGT", "acgtACGTACGTNNNNACGTACGTACGTAC"]
    counts = np.zeros((len(seqs), D), np.int16)
    for i, s in enumerate(seqs):
        b = s.encode()
        R.ref_compute(b, len(b), counts[i].ctypes.data)
    n = len(seqs)
    cos = np.zeros((n, n), np.float32)
    for i in range(n):
        for j in range(n):
            cos[i, j] = R.ref_cosine(counts[i].ctypes.data, counts[j].ctypes.data)
    # k-mer HVs (canonical pairs and palindromes included)
    kmers = ["AAAAA", "TTTTT", "ACGTA", "TACGT", "GATCC", "GGATC", "CCCCG", "CGGGG", "ATGCA"]
    khv = np.zeros((len(kmers), W), np.uint64)
    for i, km in enumerate(kmers):
        assert R.ref_kmer_hv(km.encode(), khv[i].ctypes.data) == 1
    base = np.zeros((5, W), np.uint64)
    for i, c in enumerate("ACGTN"):
        R.ref_base_hv(c.encode(), base[i].ctypes.data)
    # precompute windows + similarity (fresh cache per pair)
    refs = (clean + noisy)[:3] if fasta else [seqgen.random_seq(rng, 700) for _ in range(3)]
    refs = [r[:320] for r in refs]
    win_counts, win_rows = [], []
    for r in refs:
        out = np.zeros((200, D), np.int16)
        b = r.encode()
        nw = R.ref_precompute_windows(b, len(b), out.ctypes.data, 200)
        win_counts.append(nw)
        win_rows.append(out[:nw])
    queries = [refs[0][130:255], refs[1][40:165], seqgen.random_seq(rng, 125), refs[2][:125]]
    sim = np.zeros((len(queries), len(refs)), np.float32)
    pos = np.zeros((len(queries), len(refs)), np.int64)
    for i, q in enumerate(queries):
        for j, r in enumerate(refs):
            p = ctypes.c_size_t()
            qb, rb = q.encode(), r.encode()
            sim[i, j] = R.ref_precompute_similarity(qb, len(qb), rb, len(rb), ctypes.byref(p))
            pos[i, j] = p.value
    np.savez_compressed(
        os.path.join(HERE, "ref_d10000_k5.npz"),
        config=np.array([D, k, window, step], np.int64),
        seqs=np.array(seqs, dtype=object).astype("U"), counts=counts, cosine=cos,
        kmers=np.array(kmers).astype("U"), kmer_hv=khv, base_hv=base,
        win_refs=np.array(refs).astype("U"), win_counts=np.array(win_counts, np.int64),
        win_rows=np.concatenate(win_rows), queries=np.array(queries).astype("U"), sim=sim, pos=pos)
    print("wrote", os.path.join(HERE, "ref_d10000_k5.npz"), "seqs", n)


if __name__ == "__main__":
    main()
