"""Build step of librln_estep.so: hipcc for gfx950 (device kernels + engine), g++ for the OpenMP host modules; the
image preparation links hipFFT.

Run as ``python relion_amd/_build.py`` (by path, so the package and a stale library are not imported) or through ``__graft_entry__.build()``. Outputs go to
``relion_amd/librln_estep.so`` (objects under ``relion_amd/_build/``).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librln_estep.so")
ARCH = os.environ.get("RLN_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM_LIB = "/opt/rocm/lib"
CXX = os.environ.get("CXX_HOST", "g++")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
    # packed FP32 instructions are kept out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    for src in ("sampling.cpp", "synth.cpp", "formats.cpp", "refine_args.cpp"):
        o = os.path.join(OBJ, src.replace(".cpp", ".o"))
        _run([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in ("estep.hip", "proposal.hip", "imgprep.hip", "refprep.hip", "reconstruct.hip"):
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp, "-lgomp",
          f"-L{ROCM_LIB}", "-lhipfft", f"-Wl,-rpath,{ROCM_LIB}"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
