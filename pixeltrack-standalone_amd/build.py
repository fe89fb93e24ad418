"""Builds the package's native libraries in-tree (lib/*.so) for gfx950.

  libpxgen.so   synthetic event source (g++)
  libpxhost.so  host FED loop and the EDM-style framework with its plugins (g++)
  libpxgpu.so   HIP kernels + C ABI (hipcc --offload-arch=gfx950)
  tests/native/libpxprim_selftest.so   self-test kernels of the device primitives (tests only)

Every call compiles (no staleness check): the build is the "does it build" check.
"""
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
LIB = os.path.join(PKG, "lib")
INC = os.path.join(REPO, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             # parity: no FMA contraction (the serial reference is built without FMA)
             "-ffp-contract=off",
             # no packed FP32 instructions
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + INC]
CXX_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-pthread", "-I" + INC]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns about the packed-fp32 feature flag; keep other output
    err = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l)
    if err.strip():
        print(err, file=sys.stderr)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}")


def build_gen():
    _run(["g++"] + CXX_FLAGS + [os.path.join(PKG, "gen", "pxgen.cc"), "-o", os.path.join(LIB, "libpxgen.so")])


def build_host():
    srcs = sorted(glob.glob(os.path.join(PKG, "host", "*.cc")))
    _run(["g++"] + CXX_FLAGS + srcs + ["-o", os.path.join(LIB, "libpxhost.so"), "-ldl"])


def build_gpu():
    srcs = sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")))
    _run([HIPCC] + HIP_FLAGS + srcs + ["-o", os.path.join(LIB, "libpxgpu.so")])


def build_selftest():
    """Self-test kernels of the shared device primitives (tests/native), for the GPU tests."""
    src = os.path.join(REPO, "tests", "native", "primitives_selftest.hip")
    if os.path.exists(src):
        _run([HIPCC] + HIP_FLAGS + [src, "-o", os.path.join(REPO, "tests", "native", "libpxprim_selftest.so")])


def build_framework():
    """The job driver with the reference's module interface, linked against the C ABI."""
    srcs = sorted(glob.glob(os.path.join(PKG, "framework", "*.cc")))
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-pthread", "-I" + os.path.join(REPO, "include")] + srcs +
         ["-o", os.path.join(LIB, "pixeltrack-amd"), "-L" + LIB, "-lpxgpu", "-lpxhost",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"])


def build_all():
    os.makedirs(LIB, exist_ok=True)
    build_gen()
    build_host()
    build_gpu()
    build_framework()
    build_selftest()


if __name__ == "__main__":
    build_all()
