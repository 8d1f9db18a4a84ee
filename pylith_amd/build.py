"""Build the in-tree gfx950 shared library ``pylith_amd/libpylith_amd.so`` with hipcc.

Every HIP / C++ source under ``pylith_amd/csrc`` is compiled for gfx950 in one hipcc
invocation.  Packed-FP32 instruction selection is disabled for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpylith_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

DEVICE_FLAGS = [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
COMMON_FLAGS = ["-O3", "-fPIC", "-shared", "-std=c++17", "-Wall", "-Wno-unused-function"]


def sources():
    names = sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))
    return [os.path.join(CSRC, f) for f in names]


ROOT = os.path.dirname(HERE)
DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "integrator_driver.cpp")
DRIVER = os.path.join(ROOT, "tests", "cpp", "integrator_driver")


AMG_PROBE_SRC = os.path.join(ROOT, "tests", "cpp", "amg_probe.cpp")
AMG_PROBE = os.path.join(ROOT, "tests", "cpp", "libamg_probe.so")


POINTFN_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "pointfn_test_kernels.hip")
POINTFN_TEST = os.path.join(ROOT, "tests", "cpp", "libpointfn_test.so")


def build(verbose=False):
    cmd = [HIPCC, *COMMON_FLAGS, *DEVICE_FLAGS, "-x", "hip", *sources(), "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    build_driver(verbose)
    build_amg_probe(verbose)
    build_pointfn_test(verbose)
    return LIB


def build_pointfn_test(verbose=False):
    """Test-only pointwise kernels in the PetscPointFn form, in a library of their own that
    includes only the public header (tests/cpp/pointfn_test_kernels.hip)."""
    if not os.path.exists(POINTFN_TEST_SRC):
        return None
    cmd = [HIPCC, *COMMON_FLAGS, *DEVICE_FLAGS, "-I", os.path.join(ROOT, "include"), POINTFN_TEST_SRC,
           "-o", POINTFN_TEST]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return POINTFN_TEST


def build_amg_probe(verbose=False):
    """Host-only test shim over the multigrid set-up (g++; no device code)."""
    if not os.path.exists(AMG_PROBE_SRC):
        return None
    cmd = ["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-I", CSRC, AMG_PROBE_SRC,
           os.path.join(CSRC, "plx_amg.cpp"), "-lpthread", "-o", AMG_PROBE]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return AMG_PROBE


def build_driver(verbose=False):
    """The C++ IntegratorDomain test driver (host code over the C ABI, linked to the library)."""
    if not os.path.exists(DRIVER_SRC):
        return None
    cmd = [HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), DRIVER_SRC, "-L", HERE,
           "-lpylith_amd", "-Wl,-rpath,$ORIGIN/../../pylith_amd", "-o", DRIVER]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return DRIVER


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
