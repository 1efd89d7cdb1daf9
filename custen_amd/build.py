"""Build recipe for the custen_amd native library (gfx950).

Compiles every HIP / C++ source under custen_amd/csrc with hipcc in
relocatable-device-code mode (one code object, so device function pointers
from any translation unit can be called by the Fun kernels) and links:

  custen_amd/lib/libcusten.so   shared library exporting the C ABI of
                                include/cuSten.h
  custen_amd/lib/libcusten.a    the same objects for callers that link their
                                own device functions with hipcc -fgpu-rdc

Usage:  python -m custen_amd.build   (or custen_amd.build.build())
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
OBJ = os.path.join(ROOT, "_build")
LIBDIR = os.path.join(ROOT, "lib")
ARCH = os.environ.get("CUSTEN_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-std=c++17", "-O3", "-fPIC", "-fgpu-rdc", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unknown-pragmas",
    f"-I{os.path.join(REPO, 'include')}",
]


def _sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


HOST_ONLY = [
    "-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
    "-Wall", "-Wno-unused-result", "-Wno-unknown-pragmas",
    f"-I{os.path.join(REPO, 'include')}",
]


def _compile(src: str) -> str:
    """.hip files carry device code (HIP language, RDC); .cpp files are host
    C++ that only calls the HIP runtime API."""
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip", *COMMON, "-c", src, "-o", obj]
    else:
        cmd = [HIPCC, *HOST_ONLY, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed for {src}")
    return obj


def build(verbose: bool = False, out_dir: str | None = None, defines=()) -> str:
    """Compile and link.  `defines` (e.g. ["CUSTEN_OCC_F64_LARGE=4"]) and
    `out_dir` produce tuning variants without touching the default build."""
    global OBJ, COMMON
    lib_dir = out_dir or LIBDIR
    if defines:
        OBJ = os.path.join(lib_dir, "_obj")
        COMMON = COMMON + [f"-D{d}" for d in defines]
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(lib_dir, exist_ok=True)
    srcs = _sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    so = os.path.join(lib_dir, "libcusten.so")
    tmp = so + ".tmp"
    cmd = [HIPCC, "-fgpu-rdc", "--hip-link", f"--offload-arch={ARCH}", "-shared", *objs,
           "-o", tmp, "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(tmp, so)
    ar = os.path.join(lib_dir, "libcusten.a")
    if os.path.exists(ar):
        os.remove(ar)
    subprocess.run(["ar", "rcs", ar, *objs], check=True)
    if verbose:
        print(f"built {so}")
    return so


def build_oracle() -> str:
    odir = os.path.join(REPO, "oracle")
    subprocess.run(["make", "-s", "-C", odir, "-B"], check=True)
    return os.path.join(odir, "libcusten_oracle.so")


def build_native_callers() -> str:
    """The reference-shaped C++ callers the GPU tests run as child processes
    (tests/native/Makefile); they link against the library built above."""
    ndir = os.path.join(REPO, "tests", "native")
    subprocess.run(["make", "-s", "-C", ndir, "-B", "-j2", f"HIPCC={HIPCC}", f"ARCH={ARCH}"],
                   check=True)
    return os.path.join(ndir, "bin")


if __name__ == "__main__":
    build(verbose=True)
    build_oracle()
    build_native_callers()
