"""Copies the reference's own test data files (test/data/UFZ_CT_02_LHS.smf and
test/data/UFZ_CT_02_RHS.dvf, read by test/tensor-io.cpp:58-100) into this
folder as fixtures, and checks their SHA-256 digests.  Data only: a 539 x 539
sparse matrix (3367 entries) and its right-hand side.

    python tests/golden/copy_reference_data.py [/path/to/reference]
"""

import hashlib
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = {
    "UFZ_CT_02_LHS.smf": "4c203cd6559f0d9e7cc443d1ecfebc72a295efcd3eb5a22749c658067424d339",
    "UFZ_CT_02_RHS.dvf": "1c422fee289162eff2bd93fb791ecd718e7ffcaf2f12f62cad6850f038e3f22c",
}


def sha256(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def main(ref="/root/reference"):
    for name in FILES:
        src = os.path.join(ref, "test", "data", name)
        dst = os.path.join(HERE, name)
        shutil.copyfile(src, dst)
        digest = sha256(dst)
        assert digest == FILES[name], f"{name}: unexpected digest {digest}"
        print(name, os.path.getsize(dst), digest)


if __name__ == "__main__":
    main(*sys.argv[1:])
