"""Deterministic small taxonomy for the drop-in CLI tests.

2 phyla x 2 classes x 2 orders x 2 families x 2 genera x 2 species (64
leaves) under one kingdom, in the reference's on-disk layout: a
taxonomy_index.tsv (first line a header, which the reference skips) and
sequences_by_category/<category>.fasta.  Every category holds sequences
mutated from its parent's (600 nt), so a read's path is recoverable; reads
(250 nt) are sampled from species sequences with 1% substitutions, a few
with N runs.
"""
import gzip
import os

import numpy as np

ACGT = np.frombuffer(b"ACGT", np.uint8)
RANKS = ["p", "c", "o", "f", "g", "s"]


def _mut(rng, s, rate):
    a = np.frombuffer(s.encode(), np.uint8).copy()
    m = rng.random(len(a)) < rate
    a[m] = ACGT[rng.integers(0, 4, int(m.sum()))]
    return a.tobytes().decode()


def write(root, seed=7, n_reads=20, seq_len=600):
    rng = np.random.default_rng(seed)
    seqdir = os.path.join(root, "sequences_by_category")
    os.makedirs(seqdir, exist_ok=True)
    lines = ["seq_id\tleaf\ttaxonomy"]
    fasta = {}
    species = []

    def grow(prefix, depth, parent_seq, lineage):
        for b in range(2):
            name = f"{RANKS[depth]}__{prefix}{b}"
            s = _mut(rng, parent_seq, 0.08 - 0.01 * depth)
            fasta[name] = [s, _mut(rng, s, 0.01)]
            lin = lineage + [name]
            if depth + 1 < len(RANKS):
                grow(prefix + str(b), depth + 1, s, lin)
            else:
                species.append((name, lin))
                lines.append(f"id{len(species)}\t{name}\tk__Bacteria; " + "; ".join(lin))

    root_seq = ACGT[rng.integers(0, 4, seq_len)].tobytes().decode()
    grow("T", 0, root_seq, [])
    for name, seqs in fasta.items():
        with open(os.path.join(seqdir, name + ".fasta"), "w") as f:
            for i, s in enumerate(seqs):
                f.write(f">{name}_{i}\n{s}\n")
    with open(os.path.join(root, "taxonomy_index.tsv"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(root, "reads.fasta"), "w") as f:
        for i in range(n_reads):
            name, lin = species[int(rng.integers(0, len(species)))]
            src = fasta[name][int(rng.integers(0, 2))]
            st = int(rng.integers(0, seq_len - 250))
            r = _mut(rng, src[st:st + 250], 0.01)
            if i % 5 == 2:
                p = int(rng.integers(0, 200))
                r = r[:p] + "N" * 12 + r[p + 12:]
            f.write(f">read{i}|{name}\n{r}\n")
    return root


# ---- the cache-eviction taxonomy (tests/golden/make_evict_fixtures.py) ----

EVICT_WIDTHS = (1, 4, 8, 16, 64, 128)  # categories per rank, phylum to species


def evict_lineage(s):
    """Category names from phylum to species of species index s (0..127)."""
    return [f"{r}__E{s * w // 128}" for r, w in zip(RANKS, EVICT_WIDTHS)]


def evict_tree(root, records):
    """Pre-propagation layout: taxonomy_index.tsv (with the header line the
    reference skips) and sequences_by_category/<species>.fasta."""
    seqdir = os.path.join(root, "sequences_by_category")
    os.makedirs(seqdir, exist_ok=True)
    per = {}
    for i, s in enumerate(records):
        per.setdefault(i % 128, []).append(s)
    lines = ["seq_id\tleaf\ttaxonomy"]
    for sp in range(128):
        lin = evict_lineage(sp)
        lines.append(f"id{sp}\t{lin[-1]}\tk__Bacteria; " + "; ".join(lin))
        with open(os.path.join(seqdir, lin[-1] + ".fasta"), "w") as f:
            for j, s in enumerate(per.get(sp, [])):
                f.write(f">{lin[-1]}_{j}\n{s}\n")
    with open(os.path.join(root, "taxonomy_index.tsv"), "w") as f:
        f.write("\n".join(lines) + "\n")
    return root


def evict_records(gold):
    """The records (and short sequences) stored in gold/records.fasta.gz."""
    with gzip.open(os.path.join(gold, "records.fasta.gz"), "rt") as f:
        return _fasta_lines(f.read().splitlines())


def _fasta_lines(lines):
    out, cur = [], []
    for line in lines:
        if line.startswith(">"):
            if cur or out:
                out.append("".join(cur))
            cur = []
        elif line:
            cur.append(line)
    out.append("".join(cur))
    return out
