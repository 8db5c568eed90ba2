"""Build libqlgpu.so (HIP, gfx950) in-tree with hipcc.

    python -m ultradmrg_amd.build [--jobs N] [--debug]

Every .hip file is device code; every .cpp file is host code that only uses
the HIP runtime API.  The shared library lands next to this file.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libqlgpu.so")
OBJDIR = os.path.join(HERE, "_build")
ARCH = "gfx950"
ROCM_LIB = "/opt/rocm/lib"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          f"-I{os.path.join(ROOT, 'include')}"]
# packed FP32 instructions stay off in every HIP translation unit
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def compile_one(src: str, debug: bool) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    flags = list(COMMON)
    if debug:
        flags = [f for f in flags if f != "-O3"] + ["-O1", "-g"]
    if src.endswith(".hip"):
        cmd = [hipcc(), "-x", "hip", *flags, *DEVICE, "-c", src, "-o", obj]
    else:
        cmd = [hipcc(), *flags, "-D__HIP_PLATFORM_AMD__", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 8, debug: bool = False, verbose: bool = True) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: compile_one(s, debug), srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs,
           f"-L{ROCM_LIB}", "-lrocsolver", "-lrocblas", f"-Wl,-rpath,{ROCM_LIB}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


def build_cpp_tests(verbose: bool = True) -> list:
    """Compile the C++ programs under tests/cpp (header-only layer users) with g++."""
    tdir = os.path.join(ROOT, "tests", "cpp")
    out = []
    if not os.path.isdir(tdir):
        return out
    for f in sorted(os.listdir(tdir)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(tdir, f[:-4])
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{os.path.join(ROOT, 'include')}", os.path.join(tdir, f),
               "-o", exe, f"-L{HERE}", "-lqlgpu", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../ultradmrg_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ test build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        out.append(exe)
    if verbose:
        print(f"built {len(out)} C++ test programs", file=sys.stderr)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--debug", action="store_true")
    a = ap.parse_args()
    build(a.jobs, a.debug)
    build_cpp_tests()
