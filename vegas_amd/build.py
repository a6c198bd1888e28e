"""Build libvgx.so (HIP, gfx950) in-tree with the committed Makefile."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(force=True, jobs=8):
    args = ["make", "-C", os.path.join(HERE, "csrc"), "-j%d" % jobs]
    if force:
        args.append("-B")
    subprocess.run(args, check=True)
    return os.path.join(HERE, "libvgx.so")


if __name__ == "__main__":
    print(build())
