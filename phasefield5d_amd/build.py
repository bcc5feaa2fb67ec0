"""Build libpf5d.so in-tree with hipcc for gfx950 (no JIT, no cache outside the tree).

The library is five translation units compiled in parallel -- the host unit
(C ABI, engine, small kernels), the FFT line kernels in FP32 and in FP64, the
FFT plane kernels, and the stencil kernels -- linked into one shared object.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
UNITS = ["pf5d_capi.hip", "pf5d_kern_fft32.hip", "pf5d_kern_fft64.hip", "pf5d_kern_fftplane.hip",
         "pf5d_kern_stencil.hip", "pf5d_kern_fftreg.hip"]
OUT = os.path.join(HERE, "libpf5d.so")
OBJDIR = os.path.join(HERE, "_obj")

HIPCC_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
               # FP64 kernels keep the reference's separate roundings; FP32 kernels fuse explicitly
               "-ffp-contract=off",
               # no packed FP32 instructions
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode}) building {what}")
    return [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l and l.strip()]


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    objs = [os.path.join(OBJDIR, u.replace(".hip", ".o")) for u in UNITS]

    def compile_unit(i):
        src = os.path.join(CSRC, UNITS[i])
        return _run([hipcc()] + HIPCC_FLAGS + ["-c", "-o", objs[i], src], src)

    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        msgs = sum(ex.map(compile_unit, range(len(UNITS))), [])
    tmp = OUT + ".tmp"
    msgs += _run([hipcc()] + HIPCC_FLAGS + ["-shared", "-o", tmp] + objs, OUT)
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
