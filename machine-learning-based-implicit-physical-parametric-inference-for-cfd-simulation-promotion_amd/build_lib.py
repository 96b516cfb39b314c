"""Build libippi_lbm.so (HIP, gfx950) in-tree: <package>/lib/libippi_lbm.so.

    python <package>/build_lib.py [--debug]

Flags: -O3, gfx950 code object only, packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops), host code without FP
contraction so the host-side marker weights match the FP64 CPU arithmetic.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, "csrc", f) for f in ("ippi_lbm.hip", "ippi_train.hip", "ippi_tausearch.hip")]
OUT_DIR = os.path.join(HERE, "lib")
OUT = os.path.join(OUT_DIR, "libippi_lbm.so")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(debug=False, verbose=True):
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
           "-O1" if debug else "-O3",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-ffp-contract=off",
           "-I", os.path.join(ROOT, "include"),
           "-o", tmp] + SRC
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(debug="--debug" in sys.argv))
