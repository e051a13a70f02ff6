"""Builds the gfx950 device library libfluidsolver_amd.so in-tree with hipcc.

The device code is compiled with -ffp-contract=off (no fused multiply-add, so
FP32 results follow the reference's evaluation order), correctly rounded FP32
division and square root (hipcc's default), and without packed FP32 ops.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libfluidsolver_amd.so")
OBJDIR = os.path.join(LIBDIR, "obj")
ARCH = os.environ.get("FSA_OFFLOAD_ARCH", "gfx950")
SOURCES = ["fsa_search.hip", "fsa_resort.hip", "fsa_sph.hip", "fsa_sph2d.hip", "fsa_entities.hip", "fsa_api.hip"]
HOST_SOURCES = ["fsa_io.cpp"]  # host-only C++ (scenario files), built with g++
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-Wall", "-Wextra", "-Wno-unused-parameter"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return obj

    def compile_host(src):
        obj = os.path.join(OBJDIR, src.replace(".cpp", ".o"))
        cmd = ["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {src}:\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES) + len(HOST_SOURCES)) as ex:
        dev = [ex.submit(compile_one, s) for s in SOURCES]
        host = [ex.submit(compile_host, s) for s in HOST_SOURCES]
        objs = [f.result() for f in dev] + [f.result() for f in host]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_mathcheck()
    build_dropin()
    return LIB


MATHCHECK_LIB = os.path.join(LIBDIR, "libfsa_mathcheck.so")


def build_mathcheck():
    """Device checks of the pair-math helpers (loaded by tests/test_gpu_math.py only)."""
    tmp = MATHCHECK_LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-shared", os.path.join(CSRC, "fsa_mathcheck.hip"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for fsa_mathcheck.hip:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, MATHCHECK_LIB)
    return MATHCHECK_LIB


DROPIN_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "dropin_dam.cpp")
DROPIN_BIN = os.path.join(LIBDIR, "dropin_dam")


def build_dropin():
    """The C++ drop-in example, a plain g++ program linked against the library."""
    if not os.path.exists(DROPIN_SRC):
        return None
    inc = os.path.join(os.path.dirname(HERE), "include")
    cmd = ["g++", "-O2", "-std=c++17", "-I", inc, DROPIN_SRC, "-o", DROPIN_BIN, "-L", LIBDIR,
           "-lfluidsolver_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for the drop-in example:\n{r.stdout}\n{r.stderr}")
    return DROPIN_BIN


if __name__ == "__main__":
    print(build(verbose=True))
