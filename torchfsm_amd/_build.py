"""Build libtfsm.so (hand-written HIP for gfx950) in-tree with hipcc.

The library is compiled from torchfsm_amd/csrc/*.hip into torchfsm_amd/libtfsm.so.  Object files
go to torchfsm_amd/csrc/_build/.  No JIT cache: the built .so travels with the source tree.
"""

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
OBJDIR = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libtfsm.so")
SOURCES = ["api.hip", "k_coefs.hip", "k_transform.hip", "k_step2d.hip", "k_step2d_vc.hip", "k_step2d_conv.hip",
           "k_step2d_nsp.hip", "k_step2d_ks.hip", "k_step2d_cons.hip",
           "k_step2d_mixed.hip", "k_step3d.hip", "k_adjoint.hip",
           "k_maps.hip", "k_gfft.hip", "k_general.hip"]
ARCH = "gfx950"
# packed FP32 instructions are disabled for all device code
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: the HIP toolchain (ROCm) is required to build libtfsm.so")
    return exe


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *CFLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{res.stderr[-4000:]}")
    return obj


def build(jobs=None, verbose=False):
    """Compile every HIP source for gfx950 and link torchfsm_amd/libtfsm.so (always rebuilds)."""
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
