"""Build libglia_rd.so in-tree (hipcc for gfx950) with the Makefile in glia_amd/csrc.

    python -m glia_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=4, force=False):
    args = ["make", "-C", os.path.join(HERE, "csrc"), "-j%d" % jobs]
    if force:
        args.insert(1, "-B")
    subprocess.check_call(args)
    return os.path.join(HERE, "libglia_rd.so")


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
