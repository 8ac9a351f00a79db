"""CPU checks of the hierarchical-descent pieces (no GPU):

* lsort.h, the device restatement of std::sort's order for equal keys, against
  the host's std::sort on random, tie-heavy and adversarial inputs
  (tests/cpp/lsort_check.cpp, including inputs that reach its heapsort branch);
* the oracle's orc_descend against the reference's own identify tool, run on a
  taxonomy whose categories each hold one 100-nt sequence
  (tests/golden/tax_short, written by tests/golden/make_descent_fixtures.py):
  on it identify's beam is the descent over per-category bundles.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "bacterial-classification_amd", "python"))
GOLD = os.path.join(ROOT, "tests", "golden", "tax_short")
D_REF, K_REF = 10000, 5  # the reference build's configuration


def test_lsort_matches_std_sort(tmp_path):
    exe = str(tmp_path / "lsort_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "lsort_check.cpp")], check=True)
    p = subprocess.run([exe], capture_output=True, text=True)
    print(p.stdout)
    assert p.returncode == 0, p.stdout
    cases, heap = p.stdout.split()[1], p.stdout.split()[2]
    assert int(cases) > 400000 and int(heap.split("=")[1]) > 0


def load_tax_short():
    """categories (id order of the file), their parents, sequences; reads."""
    names, parents, seqs = [], [], []
    for line in open(os.path.join(GOLD, "categories.tsv")):
        n, p, s = line.rstrip("\n").split("\t")
        names.append(n)
        parents.append(p)
        seqs.append(s)
    idx = {n: i for i, n in enumerate(names)}
    parent = np.array([idx[p] if p != "-" else -1 for p in parents], np.int64)
    reads = []
    name = None
    for line in open(os.path.join(GOLD, "reads.fasta")):
        line = line.rstrip("\n")
        if line.startswith(">"):
            name = line[1:]
        else:
            reads.append((name, line))
    return names, parent, seqs, reads


def candidate_order(names):
    """CSR candidate lists and roots from candidate_order.tsv (categories
    not listed have no children and list themselves)."""
    idx = {n: i for i, n in enumerate(names)}
    kids = {}
    roots = []
    for line in open(os.path.join(GOLD, "candidate_order.tsv")):
        k, v = line.rstrip("\n").split("\t")
        lst = [idx[x] for x in v.split(",")]
        if k == "ROOTS":
            roots = lst
        else:
            kids[idx[k]] = lst
    off = np.zeros(len(names) + 1, np.int32)
    flat = []
    for c in range(len(names)):
        flat.extend(kids.get(c, [c]))
        off[c + 1] = len(flat)
    return off, np.array(flat, np.int32), np.array(roots, np.int32)


def expected_lines():
    lines = open(os.path.join(GOLD, "identify_expected.txt")).read().split("\n")
    return sorted(l for l in lines if ":" in l)


def format_lines(names, reads, path, score, length, exit_level):
    """identify's result lines (targets/identify.cpp:228-246): a read that
    exited early was erased from the beams and prints no path."""
    out = []
    for r, (rname, _) in enumerate(reads):
        s = rname + ":"
        if exit_level[r] < 0:
            for i in range(length[r]):
                s += f" {names[path[r, i]]} ({float(score[r, i]):g})"
        out.append(s)
    return sorted(out)


@pytest.fixture(scope="module")
def short_tree():
    import hdc_amd as H
    import oracle_py as orc
    names, parent, seqs, reads = load_tax_short()
    off, idx, roots = candidate_order(names)
    tax = H.Taxonomy(off, idx, roots, H.taxonomy_from_parents(parent).ranks)
    proto = np.stack([orc.compute(D_REF, K_REF, s)[0] for s in seqs]).astype(np.int32)
    q = np.stack([orc.compute(D_REF, K_REF, r)[0] for _, r in reads])
    return names, tax, proto, q, reads


def test_candidate_order_is_hierarchy_order():
    """candidate_order.tsv = the oracle's restatement of the reference
    Hierarchy's container order on the fixture's taxonomy_index.tsv."""
    import oracle_py as orc
    names = load_tax_short()[0]
    off, idx, roots = orc.tax_order(os.path.join(GOLD, "taxonomy_index.tsv"), names)
    o2, i2, r2 = candidate_order(names)
    assert np.array_equal(off, o2) and np.array_equal(idx, i2) and np.array_equal(roots, r2)


def test_oracle_descend_matches_reference_identify(short_tree):
    import oracle_py as orc
    names, tax, proto, q, reads = short_tree
    path, ps, ln, ex = orc.descend(proto, tax.child_off, tax.child_idx, tax.roots, q, 5, 0.0, 6)
    got = format_lines(names, reads, path, ps, ln, ex)
    assert got == expected_lines()
    assert (ex >= 0).sum() >= 3 and (ex < 0).sum() >= 50  # both outcomes are covered


def test_tax_short_results_depend_on_tie_order(short_tree):
    """The fixture exercises equal scores: with every candidate list and the
    roots reversed, some read's result changes, so matching the reference's
    output requires its candidate order and std::sort's order for ties."""
    import oracle_py as orc
    names, tax, proto, q, reads = short_tree
    a = orc.descend(proto, tax.child_off, tax.child_idx, tax.roots, q, 5, 0.0, 6)
    rev = np.concatenate([tax.child_idx[tax.child_off[c]:tax.child_off[c + 1]][::-1]
                          for c in range(tax.n_cat)])
    b = orc.descend(proto, tax.child_off, rev, tax.roots[::-1].copy(), q, 5, 0.0, 6)
    assert not np.array_equal(a[0], b[0])


def test_oracle_descend_scores_are_reference_cosine(short_tree):
    """Per-rank scores of the oracle's result paths equal the reference
    build's own cosine() (oracle/_ref) on the same bundles, bit for bit."""
    import oracle_py as orc
    if not os.path.exists(orc.REF_SO):
        pytest.skip("reference build absent (make -C oracle ref)")
    R = orc.reference()
    cfg = orc.ref_config()
    if (cfg[0], cfg[1]) != (D_REF, K_REF):
        pytest.skip("reference build at another configuration")
    names, tax, proto, q, reads = short_tree
    path, ps, ln, ex = orc.descend(proto, tax.child_off, tax.child_idx, tax.roots, q, 5, 0.0, 6)
    checked = 0
    for r in range(len(reads)):
        for i in range(ln[r]):
            p16 = proto[path[r, i]].astype(np.int16)
            assert R.ref_cosine(q[r].ctypes.data, p16.ctypes.data) == ps[r, i]
            checked += 1
    assert checked > 300
