"""Build recipe for the gfx950 engine (libopentm.so) and the FP64 oracle.

The engine is compiled with hipcc for gfx950 only, into the package
directory so the shared object travels with the source tree. The oracle
(test infrastructure) is compiled with gcc through oracle/Makefile.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libopentm.so")
SOURCES = ["otm_fine.hip", "otm_coarse.hip", "otm_vec.hip", "otm_opt.hip", "otm_mma.hip", "otm_engine.hip", "otm_batch.hip",
           "otm_framework.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build_engine(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        for src in SOURCES:
            obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
            objs.append(obj)
            cmd = [HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
            jobs.append(ex.submit(_run, cmd))
        for j in jobs:
            j.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_pybind(verbose=False):
    """The pybind11 module openTM (g++), linked against libopentm.so."""
    import sysconfig
    import pybind11
    out = os.path.join(HERE, "openTM" + sysconfig.get_config_var("EXT_SUFFIX"))
    cmd = ["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-fvisibility=hidden",
           "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
           "-I" + os.path.join(ROOT, "include"),
           os.path.join(CSRC, "openTM_pybind.cpp"), "-L" + HERE, "-lopentm",
           "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    if verbose:
        print("built", out)
    return out


DROPIN = os.path.join(ROOT, "tests", "cxx", "dropin_gpu")


def build_dropin(verbose=False):
    """The C++ drop-in driver the GPU tests run (tests/cxx/dropin_gpu.cpp):
    host code only, against include/opentm/opentm.hpp, libopentm.so and the
    HIP runtime."""
    cmd = ["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cxx", "dropin_gpu.cpp"),
           "-L" + HERE, "-lopentm", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../opentm_amd", "-Wl,-rpath,/opt/rocm/lib", "-o", DROPIN + ".tmp"]
    _run(cmd)
    os.replace(DROPIN + ".tmp", DROPIN)
    if verbose:
        print("built", DROPIN)
    return DROPIN


def build_oracle(verbose=False):
    _run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")])
    if verbose:
        print("built oracle/liboracle.so")


def build_all(verbose=False):
    build_oracle(verbose)
    build_engine(verbose)
    build_pybind(verbose)
    build_dropin(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
