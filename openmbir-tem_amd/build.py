"""Build the MI355X MBIR engine (libmbir_tem.so) for gfx950 with hipcc.

Usage: python openmbir-tem_amd/build.py   (or __graft_entry__.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmbir_tem.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-I", os.path.join(REPO, "include"),
          "-I", os.path.join(HERE, "engine")]
DEVICE = [
    "--offload-arch=gfx950",
    "-munsafe-fp-atomics",
    # no fused multiply-add contraction: the compensated sums and every kernel that
    # shares the line update then round identically, so a job's result does not
    # depend on which kernel (single job, batch, server) ran it
    "-ffp-contract=off",
    # no packed fp32 instructions: some of their forms have given wrong results on
    # this hardware without any error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
ENGINE = os.path.join(HERE, "engine")
SOURCES = [
    (os.path.join(CSRC, "mbir_kernels.hip"), ["-x", "hip"] + DEVICE),
    (os.path.join(CSRC, "mbir_capi.cpp"), ["-x", "hip"] + DEVICE),
    (os.path.join(CSRC, "mbir_serve.cpp"), ["-x", "hip"] + DEVICE),
    # the forward projection without allocation (library-wide device scratch)
    (os.path.join(CSRC, "mbir_noalloc.hip"), ["-x", "hip"] + DEVICE),
    # host adapter with the reference engine's C++ interface (no device code)
    (os.path.join(ENGINE, "HAADF_ReconstructionEngine.cpp"), ["-x", "c++"]),
    # the engine's files (MRC tilt series in, MRC volume and binary outputs out)
    (os.path.join(ENGINE, "MBIRIO.cpp"), ["-x", "c++"]),
    (os.path.join(ENGINE, "HAADF_MultiResolutionReconstruction.cpp"), ["-x", "c++"]),
    # bright-field engine and driver on the same adapter
    (os.path.join(ENGINE, "BFReconstructionEngine.cpp"), ["-x", "c++"]),
    (os.path.join(ENGINE, "BFMultiResolutionReconstruction.cpp"), ["-x", "c++"]),
    (os.path.join(ENGINE, "mbir_io_capi.cpp"), ["-x", "c++"]),
    # initial-parameter estimators (per-view reductions on the host)
    (os.path.join(ENGINE, "mbir_estimate.cpp"), ["-x", "c++"]),
]
CLI = os.path.join(HERE, "engine_cli")


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, flags in SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmd = [HIPCC] + COMMON + flags + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    cmd = [HIPCC] + COMMON + ["-x", "c++", os.path.join(ENGINE, "engine_cli.cpp"), "-I", ENGINE,
                              "-o", CLI + ".tmp", "-L", HERE, "-lmbir_tem", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(CLI + ".tmp", CLI)
    return OUT


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
