"""Build the gfx950 shared library ``libtsiir.so`` from ``csrc/``.

Explicit hipcc invocation (no torch extension machinery, no hipify): every
``.hip`` source is compiled for gfx950 into one position-independent shared
library with a plain C ABI (``include/tsiir.h``).  Packed-FP32 instruction
selection is switched off for all device code.

Usage: ``python -m torchscience_amd.build`` (or ``build()`` from Python).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_DIR = os.path.join(PKG_DIR, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libtsiir.so")
INCLUDE = os.path.join(REPO_DIR, "include")
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libtsiir.so)")


TORCH_LIB_PATH = os.path.join(LIB_DIR, "libtsiir_torch.so")
TORCH_OPS_SRC = os.path.join(CSRC, "torch_ops.cpp")


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def torch_ops_cmd(out: str) -> list[str]:
    """The PyTorch operator registration (host C++ only, no device code):
    torch's headers, HIP's runtime headers, linked against libtsiir.so
    (found next to it at run time) and torch's libraries (already loaded by
    `import torch` when the library is)."""
    import torch
    from torch.utils.cpp_extension import include_paths, library_paths
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cxx = os.environ.get("CXX", "g++")
    return [
        cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wno-unused-parameter",
        *[f"-I{p}" for p in include_paths()], "-I/opt/rocm/include", f"-I{INCLUDE}",
        "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
        TORCH_OPS_SRC, *[f"-L{p}" for p in library_paths()], f"-L{LIB_DIR}",
        "-ltsiir", "-lc10", "-lc10_hip", "-ltorch_cpu", "-ltorch_hip", "-ltorch",
        "-Wl,-rpath,$ORIGIN", "-o", out,
    ]


OBJ_DIR = os.path.join(REPO_DIR, "_build", "obj")


def compile_cmd(src: str, obj: str) -> list[str]:
    """One translation unit to a position-independent object (device code
    for gfx950 embedded)."""
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-c",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        f"-I{INCLUDE}", f"-I{CSRC}", src, "-o", obj,
    ]


def link_cmd(objs: list[str], out: str) -> list[str]:
    return [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", out]


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every source (in parallel, one hipcc per file) and link
    libtsiir.so.  Always recompiles: the build check must see the sources."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(OBJ_DIR, os.path.basename(s)[:-4] + ".o") for s in srcs]
    jobs = jobs or min(len(srcs), max(1, min(16, os.cpu_count() or 1)))

    def one(pair):
        src, obj = pair
        cmd = compile_cmd(src, obj)
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        return src, r

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(one, zip(srcs, objs)))
    failed = [(s, r) for s, r in results if r.returncode != 0]
    for s, r in results:
        # hipcc's host pass does not know the packed-FP32 feature: harmless
        err = "\n".join(ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln)
        if err.strip() and (verbose or r.returncode != 0):
            print(err, file=sys.stderr)
    if failed:
        raise RuntimeError("hipcc failed for " + ", ".join(os.path.basename(s) for s, _ in failed))
    tmp = LIB_PATH + ".tmp"
    cmd = link_cmd(objs, tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    tmp = TORCH_LIB_PATH + ".tmp"
    cmd = torch_ops_cmd(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, TORCH_LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
