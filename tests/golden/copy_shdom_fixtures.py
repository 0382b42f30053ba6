"""Copies the original-SHDOM verification files of the reference's tests
(tests/data: the polarized property file rico32x36x26w672.prp, the source
function SHDOM wrote for it, shdom_verification_source_out.out, and its
radiance output rico32x36x26w672ar.out; used by tests/test_shdom.py:175-277)
into tests/golden/shdom_rico/ as gzip files, unchanged.

Run once in a checkout that has the reference next to it:
    python tests/golden/copy_shdom_fixtures.py /root/reference
"""
import gzip
import os
import shutil
import sys

FILES = ("rico32x36x26w672.prp", "shdom_verification_source_out.out", "rico32x36x26w672ar.out")


def main(reference_root):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shdom_rico")
    os.makedirs(out, exist_ok=True)
    for name in FILES:
        src = os.path.join(reference_root, "tests", "data", name)
        with open(src, "rb") as f, gzip.GzipFile(os.path.join(out, name + ".gz"), "wb", mtime=0) as g:
            shutil.copyfileobj(f, g)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
