"""Builds mke_amd/libphelm_hip.so for gfx950 with hipcc (in-tree, no JIT).

    python -m mke_amd.build          # or __graft_entry__.build()
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libphelm_hip.so")

HIP_SOURCES = ["assemble.hip", "spmv.hip", "gmres.hip", "gmres_batch.hip", "mg.hip", "jacobian.hip", "barvortex.hip", "baroclin.hip", "block.hip"]
CPP_SOURCES = ["mesh.cpp", "capi.cpp", "serve.cpp"]

# -ffp-contract=off everywhere: the assembly must match the CPU restatement
# bit for bit, and the FP64 SpMV keeps the CSR order without fused
# multiply-add; explicit fma() calls are still fused.
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable"]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src):
    hipcc = _hipcc()
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src.endswith(".hip"):
        cmd = [hipcc, "-x", "hip", *COMMON, *DEVICE, "-c", path, "-o", obj]
    else:
        cmd = [hipcc, *COMMON, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    sources = [s for s in HIP_SOURCES + CPP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
