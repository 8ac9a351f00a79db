#!/usr/bin/env python3
"""Runs the reference's own identify and propogate-children (built from its
sources by `make -C oracle ref` into oracle/_ref, at the reference's
configuration D=10000, k=5) on the deterministic taxonomy of tests/taxdata.py
and stores their outputs under tests/golden/tax_small/ as fixtures."""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import taxdata  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")


def main():
    out = os.path.join(HERE, "tax_small")
    os.makedirs(out, exist_ok=True)
    with tempfile.TemporaryDirectory() as d:
        taxdata.write(d)
        res = os.path.join(d, "identify_out.txt")
        open(res, "w").close()
        p = subprocess.run([os.path.join(REF, "identify"), os.path.join(d, "sequences_by_category"),
                            os.path.join(d, "taxonomy_index.tsv"), os.path.join(d, "reads.fasta"),
                            res], capture_output=True, text=True, check=True)
        shutil.copy(res, os.path.join(out, "identify_expected.txt"))
        with open(os.path.join(out, "identify_stdout_expected.txt"), "w") as f:
            f.write(p.stdout)
        p = subprocess.run([os.path.join(REF, "propogate-children"),
                            os.path.join(d, "sequences_by_category"),
                            os.path.join(d, "taxonomy_index.tsv")],
                           capture_output=True, text=True, check=True)
        # digests of the tool's stdout and of every FASTA after its appends
        seqdir = os.path.join(d, "sequences_by_category")
        digest = {"stdout": hashlib.sha256(p.stdout.encode()).hexdigest(),
                  "stdout_tail": p.stdout[-2000:],
                  "fasta": {name: hashlib.sha256(open(os.path.join(seqdir, name), "rb").read())
                            .hexdigest() for name in sorted(os.listdir(seqdir))}}
        with open(os.path.join(out, "propogate_expected.json"), "w") as f:
            json.dump(digest, f, indent=1)
    print("fixtures written to", out)


if __name__ == "__main__":
    main()
