"""Builds libhipbrian.so for gfx950 in place: `python brian2_amd/build.py`.

Run as a script (it must not import the package, whose import needs the library)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # fp64 results must be bit-identical to the host restatement: no contraction
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    out = os.path.join(HERE, "libhipbrian.so")
    tmp = out + ".tmp"
    cmd = ["hipcc", *HIPCC_FLAGS, "-I", os.path.join(REPO, "include"), "-o", tmp,
           os.path.join(HERE, "csrc", "hipbrian.hip"), "-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building libhipbrian.so failed")
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
