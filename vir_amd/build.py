"""Build libvir_amd.so (HIP, gfx950) in-tree.

    python -m vir_amd.build        # or vir_amd.build.build()

Every HIP translation unit is compiled with:
  --offload-arch=gfx950        CDNA4 only, no other targets
  -ffp-contract=off            no implicit FMA: the Siddon index arithmetic
                               must round exactly as numpy's separate ops
  -Xclang -target-feature -Xclang -packed-fp32-ops
                               keep the compiler from emitting packed FP32 ops
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvir_amd.so")
OBJ = os.path.join(HERE, "_build")
SOURCES = ["vir_common.hip", "siddon.hip", "csr.hip", "sirt_ops.hip", "dd.hip", "sf.hip", "loss.hip", "geom.hip",
           "phantom.hip", "rd.hip", "sq.hip", "pdhg.hip", "fbp.hip", "tvden.hip", "sp.hip", "siddons_c.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-ffp-contract=off",
         "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
         "-I" + os.path.join(HERE, "..", "include")]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr[-6000:]}")
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
