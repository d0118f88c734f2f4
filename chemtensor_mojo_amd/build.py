"""In-tree build of libdmrg_hip.so (hipcc, gfx950) via csrc/Makefile."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")


def build(jobs: int = 8, force: bool = True, quiet: bool = True) -> str:
    """Compile every HIP source for gfx950 and link chemtensor_mojo_amd/libdmrg_hip.so."""
    args = ["make", "-C", CSRC, f"-j{jobs}"]
    if force:
        args.insert(1, "-B")
    res = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError("libdmrg_hip.so build failed:\n" + res.stdout[-8000:])
    if not quiet:
        print(res.stdout)
    return os.path.join(HERE, "libdmrg_hip.so")


if __name__ == "__main__":
    print(build(quiet=False))
