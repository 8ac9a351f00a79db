#!/usr/bin/env python3
"""Runs the reference's own identify (built from its sources by
`make -C oracle ref` into oracle/_ref, at the reference's configuration
D=10000, k=5) on a taxonomy in which every category holds ONE sequence of
100 nt.  HVCache::similarity scores a reference sequence no longer than its
125-nt window by cosine(read, whole sequence) (src/encoder.cpp:198-203), so
on this taxonomy identify's beam is exactly the descent over per-category
prototype bundles that hvk_descend and the oracle's orc_descend compute.

Tree: 3 phyla x 3 classes x 2 orders x 2 families x 2 genera x 4 species
(426 categories).  With BEAM_WIDTH 5 and 4 species per genus a level can
hold 20 candidates, more than std::sort's insertion-sort threshold.  Every
category's sequence is its parent's with substitutions; reads are 90 nt from
species sequences with 0.5-1.5% substitutions, plus random reads and reads with
N runs (which exit early or score low).

Low substitution rates leave some siblings with equal sequences, so equal
scores occur and identify's result there depends on the order in which its
unordered containers list the candidates and on std::sort's order for equal
keys.  That order is recorded too: candidate_order.tsv holds, per category
with children, its children in the order the reference's Hierarchy returns
them, and the roots in get_order("p__") order, computed by the oracle's
restatement of Hierarchy's containers (orc_tax_order) on the same
taxonomy_index.tsv.

Writes tests/golden/tax_short/: categories.tsv (name, parent, sequence),
taxonomy_index.tsv, reads.fasta, candidate_order.tsv and
identify_expected.txt (the tool's result file)."""
import os
import shutil
import subprocess
import tempfile

import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
ACGT = np.frombuffer(b"ACGT", np.uint8)
RANKS = ["p", "c", "o", "f", "g", "s"]
FANOUT = [3, 3, 2, 2, 2, 4]
RATES = [0.03, 0.025, 0.02, 0.02, 0.015, 0.015]
SEQ_LEN = 100


def _mut(rng, s, rate):
    a = np.frombuffer(s.encode(), np.uint8).copy()
    m = rng.random(len(a)) < rate
    a[m] = ACGT[rng.integers(0, 4, int(m.sum()))]
    return a.tobytes().decode()


def build(seed=11, n_reads=64):
    rng = np.random.default_rng(seed)
    cats = []  # (name, parent, seq)
    species = []

    def grow(prefix, depth, parent, parent_seq, lineage):
        for b in range(FANOUT[depth]):
            name = f"{RANKS[depth]}__X{prefix}{b}"
            s = _mut(rng, parent_seq, RATES[depth])
            cats.append((name, parent, s))
            lin = lineage + [name]
            if depth + 1 < len(RANKS):
                grow(prefix + str(b), depth + 1, name, s, lin)
            else:
                species.append((name, s, lin))

    root = ACGT[rng.integers(0, 4, SEQ_LEN)].tobytes().decode()
    grow("", 0, "-", root, [])
    reads = []
    seen = set()
    while len(reads) < n_reads:
        i = len(reads)
        if i % 16 == 5:  # random sequence
            r = ACGT[rng.integers(0, 4, 90)].tobytes().decode()
            name = f"rnd{i}"
        else:
            name_s, s, _ = species[int(rng.integers(0, len(species)))]
            st = int(rng.integers(0, SEQ_LEN - 90 + 1))
            r = _mut(rng, s[st:st + 90], float(rng.uniform(0.005, 0.015)))
            if i % 16 == 9:
                p = int(rng.integers(0, 70))
                r = r[:p] + "N" * 8 + r[p + 8:]
            name = name_s
        if r in seen:  # identify keys reads by sequence
            continue
        seen.add(r)
        reads.append((f"read{i}|{name}", r))
    return cats, species, reads


def main():
    cats, species, reads = build()
    out = os.path.join(HERE, "tax_short")
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(out, "categories.tsv"), "w") as f:
        for name, parent, s in cats:
            f.write(f"{name}\t{parent}\t{s}\n")
    with open(os.path.join(out, "reads.fasta"), "w") as f:
        for name, r in reads:
            f.write(f">{name}\n{r}\n")
    with tempfile.TemporaryDirectory() as d:
        seqdir = os.path.join(d, "sequences_by_category")
        os.makedirs(seqdir)
        for name, _, s in cats:
            with open(os.path.join(seqdir, name + ".fasta"), "w") as f:
                f.write(f">{name}_0\n{s}\n")
        with open(os.path.join(d, "taxonomy_index.tsv"), "w") as f:
            f.write("seq_id\tleaf\ttaxonomy\n")
            for i, (name, _, lin) in enumerate(species):
                f.write(f"id{i}\t{name}\tk__Bacteria; " + "; ".join(lin) + "\n")
        res = os.path.join(d, "identify_out.txt")
        open(res, "w").close()
        shutil.copy(os.path.join(out, "reads.fasta"), os.path.join(d, "reads.fasta"))
        subprocess.run([os.path.join(REF, "identify"), seqdir, os.path.join(d, "taxonomy_index.tsv"),
                        os.path.join(d, "reads.fasta"), res], capture_output=True, text=True,
                       check=True)
        shutil.copy(res, os.path.join(out, "identify_expected.txt"))
        shutil.copy(os.path.join(d, "taxonomy_index.tsv"), os.path.join(out, "taxonomy_index.tsv"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc
    names = [c[0] for c in cats]
    off, idx, roots = orc.tax_order(os.path.join(out, "taxonomy_index.tsv"), names)
    with open(os.path.join(out, "candidate_order.tsv"), "w") as f:
        f.write("ROOTS\t" + ",".join(names[i] for i in roots) + "\n")
        for c, nm in enumerate(names):
            kids = idx[off[c]:off[c + 1]]
            if len(kids) == 1 and kids[0] == c:
                continue
            f.write(nm + "\t" + ",".join(names[i] for i in kids) + "\n")
    print("fixtures written to", out)


if __name__ == "__main__":
    main()
