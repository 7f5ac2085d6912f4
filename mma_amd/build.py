"""Builds the native libraries for gfx950 with hipcc: mma_amd/lib/libsingleTDSE.so (TDSE),
mma_amd/lib/libmmahankel.so (Hankel far-field transform) and mma_amd/lib/libmmaprobe.so (the
benchmark's bandwidth probes, loaded by bench.py only).

Usage: python -m mma_amd.build   (or mma_amd.build.build() from Python)
       python -m mma_amd.build --usage            (also print each kernel's registers and spills)
       python -m mma_amd.build --sanitize DIR     (host code with AddressSanitizer and
           UndefinedBehaviorSanitizer, device code unchanged, into DIR/lib; load it with
           MMA_AMD_LIB_DIR=DIR/lib and clang's ASan runtime preloaded)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MMA_AMD_ARCH", "gfx950")

COMMON = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-I", INCLUDE, "-I", CSRC,
          # no packed FP32 instructions (v_pk_*_f32) anywhere
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

# (source, extra flags): the setup and analysis code must not contract a*b+c into FMA,
# the ground state's bitwise fixed point depends on it
SOURCES = [
    ("tdse_setup.hip", ["-ffp-contract=off"]),
    ("tdse_analysis.hip", ["-ffp-contract=off"]),
    ("tdse_propagate.hip", []),
    ("tdse_adjoint.hip", []),
    ("tdse_reverse.hip", []),
    ("ctdse_abi.cpp", ["-x", "hip"]),
    ("hankel_medium.hip", []),
    ("hankel_contract.hip", []),
    ("hankel_gemm.hip", []),
    ("hankel_abi.cpp", ["-x", "hip"]),
    ("stream_probe.hip", []),
]

LIBS = {
    "libsingleTDSE.so": ["tdse_setup.hip", "tdse_analysis.hip", "tdse_propagate.hip", "tdse_adjoint.hip",
                         "tdse_reverse.hip", "ctdse_abi.cpp"],
    "libmmahankel.so": ["hankel_medium.hip", "hankel_contract.hip", "hankel_gemm.hip", "hankel_abi.cpp"],
    # bandwidth probes for bench.py only (not part of the product)
    "libmmaprobe.so": ["stream_probe.hip"],
}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


# host-only sanitizers: each -fsanitize= applies to the host compilation alone
SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]


def build(verbose=False, resource_usage=False, sanitize=None):
    """Compile every library; with `sanitize` (a directory) build the host-sanitized variant
    there (objects in sanitize/_build, libraries in sanitize/lib) instead."""
    build_dir, lib_dir, extra_all, link_extra = BUILD, LIBDIR, [], []
    if sanitize:
        build_dir, lib_dir = os.path.join(sanitize, "_build"), os.path.join(sanitize, "lib")
        extra_all, link_extra = SANITIZE, SANITIZE + ["-shared-libsan"]
    os.makedirs(build_dir, exist_ok=True)
    os.makedirs(lib_dir, exist_ok=True)
    objs = {}
    for src, flags in SOURCES:
        obj = os.path.join(build_dir, src.rsplit(".", 1)[0] + ".o")
        extra = ["-Rpass-analysis=kernel-resource-usage"] if resource_usage and src.endswith(".hip") else []
        _run([HIPCC] + COMMON + flags + extra_all + extra + ["-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs[src] = obj
    for lib, srcs in LIBS.items():
        out = os.path.join(lib_dir, lib)
        tmp = out + ".tmp"
        _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + link_extra + ["-o", tmp] +
             [objs[s] for s in srcs], verbose)
        os.replace(tmp, out)
    return os.path.join(lib_dir, "libsingleTDSE.so")


if __name__ == "__main__":
    san = sys.argv[sys.argv.index("--sanitize") + 1] if "--sanitize" in sys.argv else None
    print(build(verbose=True, resource_usage="--usage" in sys.argv, sanitize=san))
