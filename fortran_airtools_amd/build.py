"""Build recipe of the gfx950 library (libfair_hip.so) and the Fortran drop-in modules.

hipcc compiles every csrc/*.hip for gfx950 only.  Packed-FP32 instruction selection is turned
off for all device code; the FP64 reference-path files additionally disable FMA contraction
so that their arithmetic rounds exactly as the reference's x86-64 build does.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfair_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include")]
# The FP64 reference-path files (and the grouped SIRT, whose iterates equal the single-problem
# ones bit for bit) round every product and sum separately, as the reference's x86-64 build (no
# FMA) does.  The matrix-free projector is built the same way so that its
# forward and adjoint kernels evaluate the ray geometry with identical roundings (fused
# multiply-adds are written out where wanted).
PER_FILE = {"raytrace.hip": ["-ffp-contract=off"], "sparse.hip": ["-ffp-contract=off"],
            "methods.hip": ["-ffp-contract=off"], "artwin.hip": ["-ffp-contract=off"], "group.hip": ["-ffp-contract=off"],
            "matfree.hip": ["-ffp-contract=off"], "phantoms.hip": ["-ffp-contract=off"]}


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, *PER_FILE.get(os.path.basename(src), []), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


# Host code instrumented with AddressSanitizer and UndefinedBehaviorSanitizer (the device code is
# built as usual): every sanitizer flag applies to the host pass only.
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
            "-Xarch_host", "-shared-libsan"]
SAN_DIR = os.path.join(BUILD, "san")
SAN_LIB = os.path.join(SAN_DIR, "libfair_hip_san.so")


def _compile_san(src: str) -> str:
    obj = os.path.join(SAN_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, *PER_FILE.get(os.path.basename(src), []), *SAN_HOST, "-g", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc (sanitized host) failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_hip_sanitized(verbose: bool = False) -> str:
    """_build/san/libfair_hip_san.so: the library with sanitized host code, for the tests'
    sanitizer run (the loading process preloads the clang sanitizer runtime).  Rebuilt only when a
    source is newer than it."""
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".hpp")))
    srcs.append(os.path.join(ROOT, "include", "fair_hip.h"))
    if os.path.exists(SAN_LIB) and os.path.getmtime(SAN_LIB) >= max(os.path.getmtime(f) for f in srcs):
        return SAN_LIB
    os.makedirs(SAN_DIR, exist_ok=True)
    hips = [f for f in srcs if f.endswith(".hip")]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(hips))) as ex:
        objs = list(ex.map(_compile_san, hips))
    tmp = SAN_LIB + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, SAN_LIB)
    if verbose:
        print("built", SAN_LIB)
    return SAN_LIB


def build_hip(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB




FORTRAN = os.path.join(PKG, "fortran")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the Fortran drop-in modules (fortran/*.f90, in dependency order) into
    libfair_fortran.a with their .mod files under _build/fmod."""
    order = ["fair_capi.f90", "sparse_matrices.f90", "system_generation.f90", "airms.f90", "phantomgallery.f90",
             "fair_extensions.f90"]
    srcs = [os.path.join(FORTRAN, f) for f in order if os.path.exists(os.path.join(FORTRAN, f))]
    if not srcs:
        return None
    mod = os.path.join(BUILD, "fmod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for src in srcs:
        obj = os.path.join(mod, os.path.basename(src) + ".o")
        r = subprocess.run([FLANG, "-O2", "-fPIC", "-module-dir", mod, "-c", src, "-o", obj],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed for {src}:\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    lib = os.path.join(PKG, "libfair_fortran.a")
    if os.path.exists(lib):
        os.remove(lib)
    subprocess.run(["ar", "rcs", lib, *objs], check=True)
    if verbose:
        print("built", lib)
    # example callers (reference-style programs linking the drop-in modules)
    bindir = os.path.join(PKG, "bin")
    os.makedirs(bindir, exist_ok=True)
    exdir = os.path.join(FORTRAN, "examples")
    for f in sorted(os.listdir(exdir)) if os.path.isdir(exdir) else []:
        if not f.endswith(".f90"):
            continue
        exe = os.path.join(bindir, f[:-4])
        r = subprocess.run([FLANG, "-O2", "-I", mod, os.path.join(exdir, f), lib, "-L", PKG, "-lfair_hip",
                            "-Wl,-rpath,$ORIGIN/..", "-o", exe], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed for {f}:\n{r.stdout}\n{r.stderr}")
        if verbose:
            print("built", exe)
    return lib


if __name__ == "__main__":
    build_hip(verbose=True)
    build_fortran(verbose=True)
    sys.exit(0)
