"""Build libvtem_hip.so (all HIP kernels + the C-ABI) in-tree for gfx950.

    python -m vtem_amd.build          # or vtem_amd.build.build()

Every source is compiled with hipcc --offload-arch=gfx950, packed fp32 ops disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops), -O3, and linked into one shared
library next to this file.  Objects go to vtem_amd/_build/ (git-ignored).
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvtem_hip.so")
ARCH = os.environ.get("VTEM_OFFLOAD_ARCH", "gfx950")

SOURCES = ["vt_encode.hip", "vt_gemm.hip", "vt_decode.hip", "vt_gram.hip", "vt_chol.hip",
           "vt_kpcg.hip", "vt_batch.hip", "vt_dirichlet.hip", "vt_iaf_par.hip", "vt_svd.hip",
           "vt_rnn.hip"]
LIBS = []
# per-source flags: keep the MFMA accumulators of the Cholesky's trailing update in
# ordinary VGPRs (otherwise the compiler copies them between AGPRs and VGPRs every k step)
EXTRA = {"vt_chol.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}

FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
         "-Xclang", "-packed-fp32-ops", "-fvisibility=hidden", "-Wno-unused-result",
         "-Wno-unused-command-line-argument"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the vtem HIP library cannot be built")


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *FLAGS, *EXTRA.get(src, []), "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, *LIBS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
