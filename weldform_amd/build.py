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
SOURCES = ["wf_kernels.hip", "wf_adjoint.hip", "wf_output.hip", "wf_capi.hip"]
ARCH = os.environ.get("WF_OFFLOAD_ARCH", "gfx950")
ROOT = os.path.dirname(HERE)
DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "upsetting.cpp")
DRIVER = os.path.join(LIBDIR, "upsetting")
HOOK_DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "compression_hooks.cpp")
HOOK_DRIVER = os.path.join(LIBDIR, "compression_hooks")
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


IO_SRC = os.path.join(HERE, "io", "wf_io.cpp")
IO_LIB = os.path.join(LIBDIR, "libweldform_io.so")
# the HDF5 library (C + high-level API) the XDMF writer links when present
HDF5_PREFIXES = [p for p in (os.environ.get("WF_HDF5_PREFIX"), "/opt/conda", "/usr") if p]


def hdf5_prefix():
    for p in HDF5_PREFIXES:
        inc = os.path.join(p, "include", "hdf5.h")
        for lib in ("lib", "lib64", "lib/x86_64-linux-gnu"):
            if os.path.exists(inc) and os.path.exists(os.path.join(p, lib, "libhdf5_hl.so")):
                return p, os.path.join(p, lib)
    return None, None


def build_io(verbose=False):
    """libweldform_io.so: the XDMF/HDF5 and CSV writers (host C++); built
    without the XDMF writer when no HDF5 library is found."""
    os.makedirs(LIBDIR, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-shared", "-fPIC", "-pthread", IO_SRC]
    prefix, libdir = hdf5_prefix()
    if prefix:
        cmd += ["-DWFIO_HDF5", "-I" + os.path.join(prefix, "include"), "-L" + libdir, "-lhdf5_hl", "-lhdf5",
                "-Wl,-rpath," + libdir]
    cmd += ["-o", IO_LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 or r.stderr.strip():
        sys.stderr.write(r.stdout + r.stderr)
        if r.returncode != 0:
            raise RuntimeError("g++ failed on " + IO_SRC)
    os.replace(IO_LIB + ".tmp", IO_LIB)
    return IO_LIB


def _gxx(src, out, verbose):
    """g++ build of a facade program: plain C++17, links the C-ABI only."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"), src,
           "-L" + LIBDIR, "-lweldform_hip", "-lweldform_io", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
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


def build_hook_driver(verbose=False):
    """The test driver whose boundary conditions are a host GeneralAfter
    hook (tests/cpp/compression_hooks.cpp)."""
    return _gxx(HOOK_DRIVER_SRC, HOOK_DRIVER, verbose)


def build_json_driver(verbose=False):
    """The WeldForm command-line driver: JSON scene files of the reference's
    format on the device path (weldform_amd/driver/weldform.cpp)."""
    return _gxx(JSON_DRIVER_SRC, JSON_DRIVER, verbose)


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_io(verbose=True))
    print(build_driver(verbose=True))
    print(build_hook_driver(verbose=True))
    print(build_json_driver(verbose=True))
