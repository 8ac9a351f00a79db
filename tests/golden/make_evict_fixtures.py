#!/usr/bin/env python3
"""Builds the cache-eviction taxonomy `tax_evict` from the reference's own
data file (EXAMPLE_READS.fasta: 494 full-length 16S records) and runs the
reference's own propogate-children and identify on it (oracle/_ref, built by
`make -C oracle ref` from the reference's sources at D=10000, k=5).

The taxonomy: 1 phylum > 4 classes > 8 orders > 16 families > 64 genera >
128 species; the first 448 records go 3-4 per species (record i to species
i mod 128), plus two short sequences (100 and 125 nt, cut from records) so
that similarity's whole-sequence branch (|ref| <= 125) runs.  Only species
hold sequences before propogate-children copies descendants upward, so
every 125-mer of a record appears in several categories, and the phylum
ends up with nearly every record: identify walks ~60,000 distinct windows
in that one category, past the LRU's 50,000 entries (CACHE_SIZE at
D=10000).  Which entries are evicted then depends on the exact sequence of
cache touches (reads are touched between windows), and evictions decide
which precompute chain a shared window's cached value comes from.

Stored under tests/golden/tax_evict/: the records and short sequences
(records.fasta.gz, the data of the per-species FASTA files), the taxonomy
TSV, the reads, and the reference tools' outputs: propogate-children's
stdout SHA-256 and per-FASTA SHA-256s, identify's result file and stdout.
tests/taxdata.py:evict_tree rebuilds the pre-propagation directory from these for the
tests."""
import gzip
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "tax_evict")
REF = os.path.join(ROOT, "oracle", "_ref")
sys.path.insert(0, os.path.join(ROOT, "tests"))
from taxdata import evict_tree as write_tree  # noqa: E402
FASTA = "/root/reference/EXAMPLE_READS.fasta"
N_RECORDS = 448


def read_fasta(path):
    out, cur = [], []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith(">"):
                if cur:
                    out.append("".join(cur))
                cur = []
            elif line:
                cur.append(line)
    if cur:
        out.append("".join(cur))
    return out


def main():
    recs = read_fasta(FASTA)[:N_RECORDS]
    recs += [recs[7][200:300], recs[11][400:525]]  # short: 100 and 125 nt
    rng = np.random.default_rng(5)
    reads = []
    for i in range(18):
        src = recs[int(rng.integers(0, N_RECORDS))]
        ln = int(rng.integers(150, 301))
        st = int(rng.integers(0, len(src) - ln))
        r = bytearray(src[st:st + ln].encode())
        for p in np.nonzero(rng.random(ln) < 0.01)[0]:
            r[p] = ord("ACGT"[(("ACGT".find(chr(r[p])) if chr(r[p]) in "ACGT" else 0) + 1) % 4])
        if i % 6 == 5:
            p = int(rng.integers(0, ln - 12))
            r[p:p + 12] = b"N" * 12
        reads.append(r.decode())
    os.makedirs(OUT, exist_ok=True)
    with gzip.GzipFile(os.path.join(OUT, "records.fasta.gz"), "wb", mtime=0) as f:
        f.write("".join(f">r{i}\n{s}\n" for i, s in enumerate(recs)).encode())
    with tempfile.TemporaryDirectory() as d:
        write_tree(d, recs)
        shutil.copy(os.path.join(d, "taxonomy_index.tsv"), os.path.join(OUT, "taxonomy_index.tsv"))
        seqdir = os.path.join(d, "sequences_by_category")
        p = subprocess.run([os.path.join(REF, "propogate-children"), seqdir,
                            os.path.join(d, "taxonomy_index.tsv")],
                           capture_output=True, text=True, check=True)
        # reads that are cached windows: 125 nt at step-10 offsets >= 10 of the
        # phylum's first sequences (scanned first, so their windows are the
        # first candidates for eviction); get(read) returns whichever value
        # the cache holds for that string, the sliding-window (precompute)
        # value or, after an eviction, the exact one
        with open(os.path.join(seqdir, "p__E0.fasta")) as f:
            phylum = [x for x in f.read().splitlines() if x and not x.startswith(">")]
        for j in range(10):
            src = phylum[3 * j]
            w = int(rng.integers(1, (len(src) - 126) // 10))
            reads.append(src[10 * w:10 * w + 125])
        with open(os.path.join(OUT, "reads.fasta"), "w") as f:
            for i, r in enumerate(reads):
                f.write(f">read{i}\n{r}\n")
        digest = {"stdout": hashlib.sha256(p.stdout.encode()).hexdigest(),
                  "stdout_tail": p.stdout[-2000:],
                  "fasta": {n: hashlib.sha256(open(os.path.join(seqdir, n), "rb").read())
                            .hexdigest() for n in sorted(os.listdir(seqdir))}}
        with open(os.path.join(OUT, "propogate_expected.json"), "w") as f:
            json.dump(digest, f, indent=1)
        counts = {n[:-6]: open(os.path.join(seqdir, n)).read().count(">")
                  for n in os.listdir(seqdir)}
        print("sequences per phylum after propagation:",
              {k: v for k, v in sorted(counts.items()) if k.startswith("p__")})
        res = os.path.join(d, "out.txt")
        open(res, "w").close()
        p = subprocess.run([os.path.join(REF, "identify"), seqdir,
                            os.path.join(d, "taxonomy_index.tsv"),
                            os.path.join(OUT, "reads.fasta"), res],
                           capture_output=True, text=True, check=True)
        shutil.copy(res, os.path.join(OUT, "identify_expected.txt"))
        with open(os.path.join(OUT, "identify_stdout_expected.txt"), "w") as f:
            f.write(p.stdout)
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
