"""Build libgputils_hip.so (gfx950) in-tree, plus the C++ API test binary.

    python -m gputils_amd.build            # library + tests/cpp binary
    python -m gputils_amd.build --lib-only

Every HIP source under gputils_amd/csrc is compiled with hipcc for gfx950
(in parallel), then linked into gputils_amd/lib/libgputils_hip.so.  The C++
drop-in API suite (tests/cpp/api_suite.cpp, which includes include/tensor.hpp)
is compiled with g++ and linked against that library.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gputils_amd", "csrc")
LIBDIR = os.path.join(ROOT, "gputils_amd", "lib")
OBJDIR = os.path.join(ROOT, "gputils_amd", "_build")
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "libgputils_hip.so"
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 (v_pk_*_f32) is kept out of all generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build gputils_amd")


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    return p.returncode, out


def compile_one(src: str) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    rc, out = _run([_hipcc(), *HIPCC_FLAGS, "-c", src, "-o", obj])
    if rc != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{out}")
    if out.strip():
        print(out)
    return obj


def build_lib(jobs: int | None = None) -> str:
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 16)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, LIB_NAME)
    tmp = out + ".tmp"
    rc, log = _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    if rc != 0:
        raise RuntimeError(f"link failed:\n{log}")
    os.replace(tmp, out)
    return out


def build_cpp_tests() -> str | None:
    src = os.path.join(ROOT, "tests", "cpp", "api_suite.cpp")
    if not os.path.exists(src):
        return None
    exe = os.path.join(ROOT, "tests", "cpp", "api_suite")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{INCLUDE}", src, "-o", exe,
           f"-L{LIBDIR}", "-lgputils_hip", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,$ORIGIN/../../gputils_amd/lib"]
    rc, out = _run(cmd)
    if rc != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{out}")
    return exe


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=None)
    a = ap.parse_args(argv)
    print(build_lib(a.jobs))
    if not a.lib_only:
        exe = build_cpp_tests()
        if exe:
            print(exe)
    return 0


if __name__ == "__main__":
    sys.exit(main())
