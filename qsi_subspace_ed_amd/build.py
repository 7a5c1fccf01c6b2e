"""In-tree build of the HIP library (gfx950) and of the CPU oracle.

The library is built with explicit hipcc commands into
``qsi_subspace_ed_amd/libqsi_gpu.so`` so that the built object travels with
the source tree (no JIT cache).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libqsi_gpu.so")
BUILD_DIR = os.path.join(ROOT, "_build")

SOURCES = ["qsi_gpu.hip", "qsi_block.hip", "qsi_serve.hip", "qsi_enum.hip", "qsi_solvers.cpp",
           "qsi_noalloc.hip"]
ARCH = "gfx950"

HIPCC_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (needs ROCm)")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def build_library(verbose=False):
    """Compile every HIP/C++ source for gfx950 and link libqsi_gpu.so."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([hipcc] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-o", tmp])
    os.replace(tmp, LIB)
    return LIB


CPP_CHECK_SRC = os.path.join(ROOT, "tests", "cpp", "wrapper_check.cpp")
CPP_CHECK = os.path.join(ROOT, "tests", "cpp", "wrapper_check")


FTLM_CHECK_SRC = os.path.join(ROOT, "tests", "cpp", "ftlm_check.cpp")
FTLM_CHECK = os.path.join(ROOT, "tests", "cpp", "ftlm_check")
HDF5_PREFIX = os.environ.get("QSI_HDF5_PREFIX", "/opt/conda")


def build_cpp_check():
    """Compile the C++ face's check programs against include/qsi_gpu.hpp and
    libqsi_gpu.so (found through an $ORIGIN rpath): tests/cpp/wrapper_check,
    and tests/cpp/ftlm_check (FtlmComputer::write_to_h5) when the HDF5 C
    library's headers are present."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    common = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__",
              "-I" + os.path.join(rocm, "include"), "-I" + os.path.join(ROOT, "include")]
    link = ["-L" + PKG, "-lqsi_gpu", "-L" + os.path.join(rocm, "lib"), "-lamdhip64",
            "-Wl,-rpath,$ORIGIN/../../qsi_subspace_ed_amd", "-Wl,-rpath," + os.path.join(rocm, "lib")]
    tmp = CPP_CHECK + ".tmp"
    _run(common + [CPP_CHECK_SRC] + link + ["-o", tmp])
    os.replace(tmp, CPP_CHECK)
    h5inc = os.path.join(HDF5_PREFIX, "include")
    h5lib = os.path.join(HDF5_PREFIX, "lib")
    if os.path.exists(os.path.join(h5inc, "hdf5.h")):
        tmp = FTLM_CHECK + ".tmp"
        # the system C++ runtime must be found before the HDF5 prefix's older one
        sysl = "/usr/lib/x86_64-linux-gnu"
        _run(common + ["-I" + h5inc, FTLM_CHECK_SRC] + link +
             [os.path.join(h5lib, "libhdf5.so"), "-Wl,-rpath," + sysl, "-Wl,-rpath," + h5lib,
              "-o", tmp])
        os.replace(tmp, FTLM_CHECK)
    return CPP_CHECK


def build_oracle(with_reference=None):
    """Build oracle/liboracle.so; also oracle/_ref/libref.so when the reference
    tree is present (with_reference=None: build it if it exists)."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "liboracle.so"])
    ref = os.environ.get("QSI_REFERENCE", "/root/reference")
    if with_reference is None:
        with_reference = os.path.isdir(os.path.join(ref, "components"))
    if with_reference:
        _run(["make", "-C", odir, "ref", "REF=" + ref])


if __name__ == "__main__":
    build_library(verbose=True)
    build_cpp_check()
    build_oracle()
    print("built", LIB, file=sys.stderr)
