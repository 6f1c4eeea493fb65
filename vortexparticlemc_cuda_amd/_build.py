"""Build recipe for libvpmc.so (hand-written HIP for gfx950).

Compiles every csrc/*.hip with hipcc --offload-arch=gfx950 in parallel, links
vortexparticlemc_cuda_amd/lib/libvpmc.so, and (optionally) the C++ API
driver bin/vpmc_driver.  Objects go to _build/ at the repository root.

Packed FP32 instructions are disabled (-packed-fp32-ops) for all device
code; hipcc's host pass warns that it does not know the feature.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "lib", "libvpmc.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

ARCH = os.environ.get("VPMC_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++20", "-fPIC", "-I", INC, "-I", CSRC,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=cwd)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build failed: " + os.path.basename(cmd[-1]))
    return r


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, "--offload-arch=" + ARCH] + COMMON + ["-c", src, "-o", obj]
    _run(cmd)
    return obj


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        phash_obj = ex.submit(compile_phash)   # the longest compile: start it first
        objs = list(ex.map(_compile, srcs))
        phash_obj = phash_obj.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_cxx_check()
    link_phash(phash_obj)
    if verbose:
        print("built", LIB, PHASH)
    return LIB


PHASH_SRC = os.path.join(PKG, "bindings", "phash.cpp")
PHASH = os.path.join(PKG, "lib", "phash.so")


def compile_phash():
    """the bindings module phash (pybind11 + torch C++ API, host code only):
    compiled to an object in _build/"""
    import sysconfig

    import torch
    from torch.utils import cpp_extension as ce
    inc = []
    for p in ce.include_paths() + [sysconfig.get_paths()["include"], INC, "/opt/rocm/include"]:
        inc += ["-isystem" if p.startswith(("/usr", "/opt")) else "-I", p]
    obj = os.path.join(BUILD, "phash.o")
    _run(["g++", "-O2", "-std=c++20", "-fPIC", "-fvisibility=hidden", "-c",
          "-D__HIP_PLATFORM_AMD__=1", "-DTORCH_EXTENSION_NAME=phash",
          "-DTORCH_API_INCLUDE_EXTENSION_H",
          "-D_GLIBCXX_USE_CXX11_ABI=%d" % int(torch._C._GLIBCXX_USE_CXX11_ABI)] + inc +
         [PHASH_SRC, "-o", obj])
    return obj


def link_phash(obj):
    """lib/phash.so: links libvpmc.so from the same directory and libtorch by rpath"""
    from torch.utils import cpp_extension as ce
    libdir = os.path.dirname(LIB)
    tlib = ce.library_paths()[0]
    tmp = PHASH + ".tmp"
    _run(["g++", "-shared", "-fPIC", obj, "-o", tmp, "-L", libdir, "-lvpmc", "-L", tlib,
          "-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python", "-L/opt/rocm/lib", "-lamdhip64",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + tlib, "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, PHASH)
    return PHASH


def build_phash():
    os.makedirs(BUILD, exist_ok=True)
    return link_phash(compile_phash())


CXX_CHECK_SRC = os.path.join(ROOT, "tests", "cxx", "cxx_api_check.cpp")
CXX_CHECK_BIN = os.path.join(ROOT, "tests", "cxx", "bin", "vpmc_cxx_check")


def build_cxx_check():
    """the C++ API driver used by the GPU tests (links libvpmc.so)"""
    if not os.path.exists(CXX_CHECK_SRC):
        return None
    os.makedirs(os.path.dirname(CXX_CHECK_BIN), exist_ok=True)
    libdir = os.path.dirname(LIB)
    _run([HIPCC, "--offload-arch=" + ARCH, "-O2", "-std=c++20", "-I", INC, CXX_CHECK_SRC,
          "-L", libdir, "-lvpmc", "-Wl,-rpath," + libdir, "-Wl,-rpath,$ORIGIN/../../../"
          "vortexparticlemc_cuda_amd/lib", "-o", CXX_CHECK_BIN])
    return CXX_CHECK_BIN


if __name__ == "__main__":
    build(verbose=True)
