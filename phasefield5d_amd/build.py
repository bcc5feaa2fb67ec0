"""Build libpf5d.so in-tree with hipcc for gfx950 (no JIT, no cache outside the tree)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "pf5d_capi.hip")
OUT = os.path.join(HERE, "libpf5d.so")

HIPCC_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
               # FP64 kernels keep the reference's separate roundings; FP32 kernels fuse explicitly
               "-ffp-contract=off",
               # no packed FP32 instructions
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(verbose=False):
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode}) building {SRC}")
    if verbose:
        msgs = [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if msgs:
            print("\n".join(msgs))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
