"""Builds the gfx950 circuit library in-tree with hipcc: ``libckt.so``, and
its bounds-checked (``libckt_checked.so``) and sanitized
(``libckt_san.so``) variants.

Every HIP and host source is compiled for gfx950 only, with
``-ffp-contract=off`` (explicit fmaf() is the only fused arithmetic) and with
packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libckt.so")
DRIVER = os.path.join(HERE, "ckt_sim")
BUILD_DIR = os.path.join(HERE, "_build")
ARCH = "gfx950"

SOURCES = ["ckt_kernels.hip", "ckt_devgen.hip", "ckt_common.cpp", "ckt_generator.cpp",
           "ckt_engine.cpp", "ckt_ctx.cpp", "ckt_batch.cpp", "ckt_server.cpp"]

COMMON_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
                "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-I" + os.path.join(HERE, "..", "include")]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


# Build variants: (library, extra compile flags, extra link flags).
#   release  the product;
#   checked  with the kernels' bounds checks (CKT_BOUNDS_CHECKS);
#   san      bounds checks, and the host code under AddressSanitizer and
#            UndefinedBehaviorSanitizer (load it with the sanitizer runtime
#            preloaded; the device code is not instrumented).
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-shared-libsan"]
VARIANTS = {
    "release": (LIB, [], []),
    "checked": (os.path.join(HERE, "libckt_checked.so"), ["-DCKT_BOUNDS_CHECKS"], []),
    "san": (os.path.join(HERE, "libckt_san.so"),
            ["-DCKT_BOUNDS_CHECKS", "-g", "-fno-omit-frame-pointer", *SAN_HOST], SAN_HOST),
}


def sanitizer_runtime() -> str:
    """The shared AddressSanitizer runtime the `san` variant links against
    (to preload into a host process that loads it)."""
    import glob
    base = os.path.dirname(os.path.dirname(os.path.realpath(hipcc())))
    hits = sorted(glob.glob(os.path.join(base, "lib", "llvm", "lib", "clang", "*", "lib", "linux",
                                         "libclang_rt.asan-x86_64.so")))
    if not hits:
        raise RuntimeError("libclang_rt.asan-x86_64.so not found under the ROCm LLVM tree")
    return hits[-1]


def _compile(job) -> str:
    variant, src = job
    flags = VARIANTS[variant][1]
    out_dir = os.path.join(BUILD_DIR, variant)
    obj = os.path.join(out_dir, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), "--offload-arch=" + ARCH, *COMMON_FLAGS, *flags, "-x", "hip", "-c",
           os.path.join(CSRC, src), "-o", obj]
    if variant == "san":
        cmd = [c if c != "-O3" else "-O1" for c in cmd]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def _link(variant, objs):
    lib, _, link_flags = VARIANTS[variant]
    tmp = lib + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", *link_flags, "-o", tmp, *objs,
           "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)


def build(verbose: bool = False, variants=("release", "checked", "san")) -> str:
    for v in variants:
        os.makedirs(os.path.join(BUILD_DIR, v), exist_ok=True)
    jobs = [(v, src) for v in variants for src in SOURCES]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, jobs))
    for v in variants:
        _link(v, [o for (vv, _), o in zip(jobs, objs) if vv == v])
    # Host driver (region_main replacement) linked against the library.
    cmd = [hipcc(), "-O2", "-std=c++17", "-o", DRIVER + ".tmp", os.path.join(CSRC, "ckt_main.cpp"),
           "-L" + HERE, "-lckt", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print("built", ", ".join(VARIANTS[v][0] for v in variants), "and", DRIVER)
    return LIB


if __name__ == "__main__":
    build(verbose=True, variants=tuple(sys.argv[1:]) or ("release", "checked", "san"))
    sys.exit(0)
