"""Build libfans_gpu.so (HIP, gfx950) in-tree.

    python -m fans_amd.build            # compile the HIP library
    python -m fans_amd.build --verbose

The library is a single translation unit (csrc/fans_gpu.hip including the
kernel headers) compiled with hipcc for gfx950 only.  Packed FP32 VALU ops are
disabled (-packed-fp32-ops) so the compiler never emits v_pk_{add,mul,fma}_f32.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfans_gpu.so")
ARCH = os.environ.get("FANS_GPU_ARCH", "gfx950")


def hipcc():
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if p and os.path.exists(p):
            return p
    raise RuntimeError("hipcc not found")


def flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-ffp-contract=fast", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
        "-Wno-unused-but-set-variable", "-Wno-unknown-pragmas",
        f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
    ]


def build(verbose=False):
    src = os.path.join(CSRC, "fans_gpu.hip")
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + flags() + ["-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed to build libfans_gpu.so")
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)
    os.replace(tmp, LIB)
    return LIB


HOST_SRC = [os.path.join(HERE, "host", f) for f in ("main.cpp", "reader.cpp", "h5dyn.cpp")]
CLI = os.path.join(HERE, "FANS")
HDF5_PREFIX = os.environ.get("FANS_HDF5_PREFIX", "/opt/conda")


def build_cli(verbose=False):
    """The drop-in CLI `fans_amd/FANS in.json out.h5`: C++ host (g++) over the C
    ABI; the serial HDF5 headers come from FANS_HDF5_PREFIX (default
    /opt/conda), the library is loaded at run time (host/h5dyn.cpp)."""
    if not os.path.exists(LIB):
        build(verbose)
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{os.path.join(HDF5_PREFIX, 'include')}", "-o", CLI + ".tmp"] + HOST_SRC + [
        f"-L{HERE}", "-lfans_gpu", "-Wl,-rpath,$ORIGIN", "-ldl"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed to build the FANS CLI")
    os.replace(CLI + ".tmp", CLI)
    return CLI


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
    print(build_cli(verbose="--verbose" in sys.argv))
