"""Build libg5hip.so (gfx950) in-tree with hipcc.

    python -m gpugraph500_amd.build        # or build() from Python

Every .hip file under csrc/ is compiled for gfx950 and linked into
gpugraph500_amd/libg5hip.so next to this file.  The C++ host side under host/
is built as well (drivers g500_hip and g500_serve).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libg5hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("G5_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(OBJ, os.path.basename(s)[:-4] + ".o") for s in srcs]

    def compile_one(pair):
        src, obj = pair
        out = _run([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
        if verbose and out.strip():
            print(out)
        return obj

    with ThreadPoolExecutor(max_workers=jobs or min(8, len(srcs))) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_host()
    return LIB


HOST = os.path.join(HERE, "host")
DRIVER = os.path.join(HERE, "g500_hip")
SERVE = os.path.join(HERE, "g500_serve")
HOST_FLAGS = ["-O2", "-std=c++17", "-Wall", "-Wextra", "-pthread", "-I" + os.path.join(ROOT, "include"), "-I" + HOST]


def host_sources():
    """The host library sources (everything under host/ but the drivers)."""
    return sorted(os.path.join(HOST, f) for f in os.listdir(HOST)
                  if f.endswith(".cpp") and not f.startswith("g500_"))


def drivers():
    """host/g500_<name>.cpp -> gpugraph500_amd/g500_<name>"""
    return {os.path.join(HERE, f[:-4]): os.path.join(HOST, f) for f in sorted(os.listdir(HOST))
            if f.startswith("g500_") and f.endswith(".cpp")}


TESTS = os.path.join(ROOT, "tests")
HIPSIMD_TEST = os.path.join(TESTS, "bin", "hipsimd_gpu")


def build_host():
    """The C++ host side (HipBFS, the emulated 2D grid, the serving engine and
    the g500_* drivers): plain g++ against the C ABI of libg5hip.so, found
    next to the binaries.  Also the GPU tests' C++ driver of the codec schema
    (tests/hipsimd_gpu.cpp -> tests/bin/hipsimd_gpu), so that the tests run
    on a GPU box without building anything there."""
    cxx = os.environ.get("CXX", "g++")
    for exe, main in drivers().items():
        tmp = exe + ".tmp"
        _run([cxx] + HOST_FLAGS + host_sources() + [main] + ["-o", tmp, "-L" + HERE, "-lg5hip", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, exe)
    src = os.path.join(TESTS, "hipsimd_gpu.cpp")
    if os.path.exists(src):
        os.makedirs(os.path.dirname(HIPSIMD_TEST), exist_ok=True)
        tmp = HIPSIMD_TEST + ".tmp"
        _run([cxx] + HOST_FLAGS + [src, "-o", tmp, "-L" + HERE, "-lg5hip", "-Wl,-rpath,$ORIGIN/../../gpugraph500_amd"])
        os.replace(tmp, HIPSIMD_TEST)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
