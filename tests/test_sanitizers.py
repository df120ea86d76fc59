"""CPU sanitizer builds of the oracle and of the C-ABI's host code.

* The oracle (C + OpenMP) with AddressSanitizer and UndefinedBehaviorSanitizer
  (gcc), and with ThreadSanitizer (clang with the LLVM OpenMP runtime and its
  Archer tool, which tells the race detector about OpenMP's barriers; with
  GCC's libgomp every implicit barrier reads as a race): tests/cpp/
  oracle_sanitize.c runs Leapfrog with damage, heat conduction and each
  contact law, and KickDrift with the kernel-gradient correction, on two
  threads.
* The C-ABI library rebuilt with its host code instrumented (address and
  undefined behaviour; the device code is not), driven by tests/cpp/
  capi_args.c: null and invalid arguments to every entry point give error
  codes, the cylinder generator stays within the buffer it is given.
* The C++ facade's host program (tests/cpp/facade_host.cpp) under the same
  sanitizers.

No GPU is used: with no device, wf_create validates its arguments and then
returns WF_EHIP."""
import os
import shutil
import subprocess

import pytest

from weldform_amd import build as wfbuild

ROOT = wfbuild.ROOT
CLANG = "/opt/rocm/llvm/bin/clang"
ARCHER = "/opt/rocm/llvm/lib/libarcher.so"
SAN_MSGS = ("runtime error", "AddressSanitizer", "LeakSanitizer", "ThreadSanitizer")


def _run(cmd, env=None, timeout=600):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)
    return r.returncode, r.stdout + r.stderr


def _clean(out):
    return not any(m in out for m in SAN_MSGS)


def test_oracle_address_and_undefined(tmp_path):
    exe = str(tmp_path / "oracle_asan")
    rc, out = _run(["gcc", "-std=c11", "-D_GNU_SOURCE", "-O1", "-g", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fopenmp",
                    "-ffp-contract=off", "-I" + os.path.join(ROOT, "oracle"),
                    os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.c"), os.path.join(ROOT, "oracle", "wf_oracle.c"),
                    "-lm", "-o", exe])
    assert rc == 0, out
    rc, out = _run([exe])
    print(out[-600:])
    assert rc == 0 and _clean(out) and "0 failures" in out, out[-3000:]


@pytest.mark.skipif(not (os.path.exists(CLANG) and os.path.exists(ARCHER)), reason="no clang with Archer")
def test_oracle_thread_sanitizer(tmp_path):
    exe = str(tmp_path / "oracle_tsan")
    libdir = os.path.dirname(ARCHER)
    rc, out = _run([CLANG, "-std=c11", "-D_GNU_SOURCE", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=thread",
                    "-fopenmp", "-ffp-contract=off", "-I" + os.path.join(ROOT, "oracle"),
                    os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.c"), os.path.join(ROOT, "oracle", "wf_oracle.c"),
                    "-lm", "-Wl,-rpath," + libdir, "-o", exe])
    assert rc == 0, out
    env = dict(os.environ, OMP_TOOL_LIBRARIES=ARCHER, TSAN_OPTIONS="halt_on_error=1")
    rc, out = _run([exe], env=env, timeout=900)
    print(out[-600:])
    assert rc == 0 and _clean(out) and "0 failures" in out, out[-3000:]


@pytest.fixture(scope="module")
def sanitized_lib(tmp_path_factory):
    """libweldform_hip built with the host code instrumented."""
    d = tmp_path_factory.mktemp("asanlib")
    objs, procs = [], []
    for src in wfbuild.SOURCES:
        obj = str(d / src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [wfbuild.hipcc(), "--offload-arch=" + wfbuild.ARCH, "-std=c++17", "-fPIC", "-O1", "-g",
               "-fno-omit-frame-pointer", "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-c", os.path.join(wfbuild.CSRC, src), "-o", obj]
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
    for p in procs:
        out, _ = p.communicate(timeout=900)
        assert p.returncode == 0, out[-3000:]
    lib = str(d / "libweldform_hip_asan.so")
    rc, out = _run([wfbuild.hipcc(), "--offload-arch=" + wfbuild.ARCH, "-shared", "-fPIC", "-Xarch_host",
                    "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", *objs, "-o", lib])
    assert rc == 0, out
    return d


def test_capi_argument_checks(sanitized_lib, tmp_path):
    exe = str(tmp_path / "capi_args")
    rc, out = _run([CLANG, "-std=c11", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "capi_args.c"), "-L" + str(sanitized_lib),
                    "-lweldform_hip_asan", "-Wl,-rpath," + str(sanitized_lib), "-Wl,-rpath,/opt/rocm/lib",
                    "-o", exe])
    assert rc == 0, out
    # the HIP runtime keeps its allocations until exit: leak checking off
    rc, out = _run([exe], env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    print(out[-600:])
    assert rc == 0 and _clean(out) and "0 failures" in out, out[-3000:]


def test_facade_host_program(sanitized_lib, tmp_path):
    if not os.path.exists(wfbuild.IO_LIB):
        wfbuild.build_io()
    # the facade links libweldform_hip by name: point it at the instrumented copy
    shutil.copy(str(sanitized_lib / "libweldform_hip_asan.so"), str(tmp_path / "libweldform_hip.so"))
    exe = str(tmp_path / "facade_host")
    rc, out = _run([CLANG + "++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "facade_host.cpp"),
                    "-L" + str(tmp_path), "-L" + wfbuild.LIBDIR, "-lweldform_hip", "-lweldform_io",
                    "-Wl,-rpath," + str(tmp_path), "-Wl,-rpath," + wfbuild.LIBDIR, "-Wl,-rpath,/opt/rocm/lib",
                    "-o", exe])
    assert rc == 0, out
    rc, out = _run([exe], env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    assert rc == 0 and _clean(out), out[-3000:]
    assert out.count("fatal ") >= 8
