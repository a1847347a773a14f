"""Builds liblensim.so for gfx950 in-tree (jaxtronomy_amd/_lib/), and its
bounds-checked debug form liblensim_checked.so (-DLENSIM_BOUNDS_CHECK: every
kernel access to a caller or workspace buffer is tested against the call's
buffers; selected by LENSIM_CHECKED=1 or backend.use_checked()).

hipcc cross-compiles without a GPU.  Packed-FP32 instructions are disabled on
purpose (-target-feature -packed-fp32-ops); the host pass warns that it does
not know that feature, which is harmless.  Matrix-core results stay in vector
registers (-amdgpu-mfma-vgpr-form): with accumulation registers the compiler
copied the float64 convolution's accumulators between the two register files
on every loop iteration.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "lensim.hip")
OUT = os.path.join(HERE, "_lib", "liblensim.so")
OUT_CHECKED = os.path.join(HERE, "_lib", "liblensim_checked.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    "-I" + os.path.join(ROOT, "include"),
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def build_library(verbose=False, checked=False):
    out_path = OUT_CHECKED if checked else OUT
    os.makedirs(os.path.dirname(out_path), exist_ok=True)
    tmp = out_path + ".tmp%d" % os.getpid()
    cmd = [hipcc()] + HIPCC_FLAGS + (["-DLENSIM_BOUNDS_CHECK"] if checked else []) + ["-o", tmp, SRC]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise RuntimeError("hipcc failed (%d):\n%s" % (res.returncode, out))
    if verbose and out.strip():
        print(out)
    os.replace(tmp, out_path)  # atomic: a process that already loaded the old .so keeps it
    return out_path


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_library(verbose=True, checked=True))
    sys.exit(0)
