"""Build libweldform_hip.so for gfx950 with hipcc (in-tree, no JIT), and the
C++ upsetting driver that runs through the SPH::Domain facade.

    python -m weldform_amd.build          # or weldform_amd.build.build()
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libweldform_hip.so")
SOURCES = ["wf_kernels.hip", "wf_adjoint.hip", "wf_capi.hip"]
ARCH = os.environ.get("WF_OFFLOAD_ARCH", "gfx950")
ROOT = os.path.dirname(HERE)
DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "upsetting.cpp")
DRIVER = os.path.join(LIBDIR, "upsetting")
JSON_DRIVER_SRC = os.path.join(HERE, "driver", "weldform.cpp")
JSON_DRIVER = os.path.join(LIBDIR, "WeldForm")

HIPCC_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 instructions stay off; hipcc's host pass warns that it does
    # not know the feature, which is harmless
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, src.replace(".hip", ".o"))
        cmd = [hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed on " + src)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def _gxx(src, out, verbose):
    """g++ build of a facade program: plain C++17, links the C-ABI only."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"), src,
           "-L" + LIBDIR, "-lweldform_hip", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 or r.stderr.strip():
        sys.stderr.write(r.stdout + r.stderr)
        if r.returncode != 0:
            raise RuntimeError("g++ failed on " + src)
    os.replace(out + ".tmp", out)
    return out


def build_driver(verbose=False):
    """The C++ upsetting driver of the tests (tests/cpp/upsetting.cpp)."""
    return _gxx(DRIVER_SRC, DRIVER, verbose)


def build_json_driver(verbose=False):
    """The WeldForm command-line driver: JSON scene files of the reference's
    format on the device path (weldform_amd/driver/weldform.cpp)."""
    return _gxx(JSON_DRIVER_SRC, JSON_DRIVER, verbose)


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_driver(verbose=True))
