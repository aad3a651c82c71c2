"""Builds libflagsparse_hip.so (gfx950) in-tree with hipcc.

Usage: python -m flagsparse_public_amd.build [--jobs N] [--force]

Each ``csrc/*.hip`` file is compiled to an object in ``csrc/_build/`` (in
parallel), then linked into ``flagsparse_public_amd/libflagsparse_hip.so``.
Packed-fp32 instructions are disabled for all device code
(``-target-feature -packed-fp32-ops``).

``csrc/vendor_rocsparse.cpp`` (host code only) is linked on its own into
``libfs_vendor.so`` against rocSPARSE: the vendor baseline the benchmark
harness times beside the library; the product library does not link it.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
BUILD_DIR = os.path.join(CSRC, "_build")
LIB_NAME = "libflagsparse_hip.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)
ARCH = "gfx950"
VENDOR_LIB_PATH = os.path.join(PKG_DIR, "libfs_vendor.so")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")

COMPILE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-mcode-object-version=5",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(PKG_DIR, '..', 'include')}",
]
LINK_FLAGS = [
    f"--offload-arch={ARCH}",
    "-shared",
    "-fPIC",
    f"-L{os.path.join(ROCM, 'lib')}",
    "-lamdhip64",
    f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}",
]


HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{os.path.join(PKG_DIR, '..', 'include')}"]
# host-only C++ sources of the main library (vendor_rocsparse.cpp builds its own)
HOST_SOURCES = ("mtx_io.cpp",)


def _compile(src: str, obj: str) -> tuple[str, int, str]:
    if src.endswith(".cpp"):
        cmd = ["g++", *HOST_FLAGS, "-c", src, "-o", obj]
    else:
        cmd = [HIPCC, *COMPILE_FLAGS, "-c", src, "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    return src, proc.returncode, proc.stdout + proc.stderr


def build(jobs: int | None = None, force: bool = True, verbose: bool = False) -> str:
    """Compile every HIP source and link the shared library; returns its path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip"))) + [
        os.path.join(CSRC, f) for f in HOST_SOURCES]
    headers = sorted(glob.glob(os.path.join(CSRC, "*.hpp"))) + [
        os.path.join(PKG_DIR, "..", "include", "flagsparse_hip.h")]
    newest_header = max(os.path.getmtime(h) for h in headers)
    todo = []
    objs = []
    for src in sources:
        obj = os.path.join(BUILD_DIR, os.path.splitext(os.path.basename(src))[0] + ".o")
        objs.append(obj)
        stale = (force or not os.path.exists(obj)
                 or os.path.getmtime(obj) < max(os.path.getmtime(src), newest_header))
        if stale:
            todo.append((src, obj))
    jobs = jobs or min(len(todo) or 1, os.cpu_count() or 4, 16)
    failures = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as pool:
        for src, rc, log in pool.map(lambda so: _compile(*so), todo):
            if rc != 0:
                failures.append((src, log))
            elif verbose and log.strip():
                print(log, file=sys.stderr)
    if failures:
        for src, log in failures:
            print(f"--- hipcc failed on {src}\n{log}", file=sys.stderr)
        raise RuntimeError(f"hipcc failed on {len(failures)} source file(s)")
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, *objs, *LINK_FLAGS, "-o", tmp]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}{proc.stderr}")
    os.replace(tmp, LIB_PATH)
    build_vendor()
    return LIB_PATH


def build_vendor() -> str:
    """Host-only C++ wrapper over rocSPARSE (benchmark baseline)."""
    src = os.path.join(CSRC, "vendor_rocsparse.cpp")
    tmp = VENDOR_LIB_PATH + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__",
           f"-I{os.path.join(ROCM, 'include')}", f"-I{os.path.join(PKG_DIR, '..', 'include')}",
           src, "-o", tmp, f"-L{os.path.join(ROCM, 'lib')}", "-lrocsparse", "-lamdhip64",
           f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"vendor baseline build failed:\n{proc.stdout}{proc.stderr}")
    os.replace(tmp, VENDOR_LIB_PATH)
    return VENDOR_LIB_PATH


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--incremental", action="store_true",
                    help="only recompile sources newer than their objects")
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args(argv)
    path = build(jobs=args.jobs, force=not args.incremental, verbose=args.verbose)
    print(path)
    return 0


if __name__ == "__main__":
    sys.exit(main())
