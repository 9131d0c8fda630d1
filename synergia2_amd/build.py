"""Build recipe of the native libraries (hipcc, gfx950), run in-tree.

    python -m synergia2_amd.build          # libsc3d_hip.so + libsc3d_host.so

The objects and shared libraries are written next to their sources so that
they travel with the tree; they are git-ignored.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SC3D_ARCH", "gfx950")

HIP_SOURCES = ["sc3d_fft.hip", "sc3d_fft_any.hip", "sc3d_particles.hip", "sc3d_maps.hip", "sc3d_capi.hip",
               "sc3d_serve.hip", "sc3d_quirk.hip", "sc3d_noalloc.hip"]
LIB = os.path.join(HERE, "libsc3d_hip.so")
HOST_LIB = os.path.join(HERE, "libsc3d_host.so")

# -ffp-contract=off: no FMA contraction anywhere (bit-exact cell indices and
# weights, same arithmetic as the CPU restatement).  Packed FP32 ops are
# disabled for the device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r


def build_hip(jobs=4):
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB, *objs])
    return LIB


def build_host():
    """C++ operator class (reference API mirror) over the C ABI."""
    srcs = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST)) if f.endswith(".cc")]
    if not srcs:
        return None
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__",
          "-I", os.path.join(ROOT, "include"), "-I", HOST, "-I", "/opt/rocm/include",
          *srcs, "-o", HOST_LIB, "-L", HERE, "-lsc3d_hip", "-L", "/opt/rocm/lib",
          "-lamdhip64", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,/opt/rocm/lib"])
    return HOST_LIB


DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "operator_driver.cc")
DRIVER = os.path.join(ROOT, "tests", "cpp", "operator_driver")


def build_driver():
    """The C++ test driver of the operator class (used by a GPU test)."""
    _run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", HOST, "-I",
          os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", DRIVER_SRC, "-o", DRIVER,
          "-L", HERE, "-lsc3d_host", "-lsc3d_hip", "-L", "/opt/rocm/lib", "-lamdhip64",
          f"-Wl,-rpath,{HERE}", "-Wl,-rpath,/opt/rocm/lib"])
    return DRIVER


def build_all():
    build_hip()
    build_host()
    build_driver()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
