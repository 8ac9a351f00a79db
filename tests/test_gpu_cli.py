"""The drop-in command-line tools (bacterial-classification_amd/bin) on the
GPU reproduce the reference's own identify / propogate-children (fixtures
written by tests/golden/make_cli_fixtures.py from the reference's build) on a
deterministic taxonomy: byte-identical result file and logs, identical FASTA
appends."""
import hashlib
import json
import os
import subprocess

import pytest

import taxdata

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bacterial-classification_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden", "tax_small")


def test_identify_matches_reference(tmp_path):
    d = taxdata.write(str(tmp_path))
    res = os.path.join(d, "out.txt")
    open(res, "w").close()
    print("identify on", d, flush=True)
    p = subprocess.run([os.path.join(BIN, "identify"), os.path.join(d, "sequences_by_category"),
                        os.path.join(d, "taxonomy_index.tsv"), os.path.join(d, "reads.fasta"), res],
                       capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr[-2000:]
    assert open(res).read() == open(os.path.join(GOLD, "identify_expected.txt")).read()
    assert p.stdout == open(os.path.join(GOLD, "identify_stdout_expected.txt")).read()


def test_propogate_children_matches_reference(tmp_path):
    d = taxdata.write(str(tmp_path))
    print("propogate-children on", d, flush=True)
    p = subprocess.run([os.path.join(BIN, "propogate-children"),
                        os.path.join(d, "sequences_by_category"),
                        os.path.join(d, "taxonomy_index.tsv")],
                       capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr[-2000:]
    want = json.load(open(os.path.join(GOLD, "propogate_expected.json")))
    assert p.stdout[-2000:] == want["stdout_tail"]
    assert hashlib.sha256(p.stdout.encode()).hexdigest() == want["stdout"]
    seqdir = os.path.join(d, "sequences_by_category")
    got = {n: hashlib.sha256(open(os.path.join(seqdir, n), "rb").read()).hexdigest()
           for n in sorted(os.listdir(seqdir))}
    assert got == want["fasta"]


# ---- past the LRU's capacity: the cache-eviction taxonomy ----
EVICT = os.path.join(ROOT, "tests", "golden", "tax_evict")


@pytest.fixture(scope="module")
def evict_propagated(tmp_path_factory):
    """The 16S taxonomy of tests/golden/tax_evict after the drop-in
    propogate-children (whose appends must equal the reference's)."""
    d = str(tmp_path_factory.mktemp("evict"))
    taxdata.evict_tree(d, taxdata.evict_records(EVICT))
    print("propogate-children on", d, flush=True)
    p = subprocess.run([os.path.join(BIN, "propogate-children"),
                        os.path.join(d, "sequences_by_category"),
                        os.path.join(d, "taxonomy_index.tsv")],
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-2000:]
    return d, p.stdout


def test_propogate_children_16s_matches_reference(evict_propagated):
    d, stdout = evict_propagated
    want = json.load(open(os.path.join(EVICT, "propogate_expected.json")))
    assert stdout[-2000:] == want["stdout_tail"]
    assert hashlib.sha256(stdout.encode()).hexdigest() == want["stdout"]
    seqdir = os.path.join(d, "sequences_by_category")
    got = {n: hashlib.sha256(open(os.path.join(seqdir, n), "rb").read()).hexdigest()
           for n in sorted(os.listdir(seqdir))}
    assert got == want["fasta"]


def test_identify_past_cache_capacity_matches_reference(evict_propagated):
    """~60,000 distinct windows at level 0 against the 50,000-entry LRU: the
    drop-in must repeat the reference's cache touches (get(window) then
    get(read) per window, get(ref) before get(read) for short refs) so that
    the same entries are evicted and shared windows keep the same values."""
    d, _ = evict_propagated
    res = os.path.join(d, "out.txt")
    open(res, "w").close()
    print("identify (cache eviction) on", d, flush=True)
    p = subprocess.run([os.path.join(BIN, "identify"), os.path.join(d, "sequences_by_category"),
                        os.path.join(d, "taxonomy_index.tsv"), os.path.join(EVICT, "reads.fasta"),
                        res], capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-2000:]
    assert open(res).read() == open(os.path.join(EVICT, "identify_expected.txt")).read()
    assert p.stdout == open(os.path.join(EVICT, "identify_stdout_expected.txt")).read()
