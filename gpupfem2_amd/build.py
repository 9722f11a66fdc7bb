"""Build recipe of libpfem2_hip.so (hipcc, gfx950) and of the CPU oracle.

    python -m gpupfem2_amd.build          # library + C++ examples + oracle
    python -m gpupfem2_amd.build --lib    # library + C++ examples

Every .hip file under csrc/ is compiled separately (in parallel) into csrc/../_build/*.o and linked
into gpupfem2_amd/libpfem2_hip.so. -ffp-contract=off keeps every floating-point expression rounded
as written, which is what makes the kernels bit-comparable with a sequential CPU evaluation.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpfem2_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PFEM2_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-I" + os.path.join(ROOT, "include"),
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass does not know the packed-fp32 feature and says so; that warning is harmless
    err = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return src, obj, p.returncode, err


def build_lib(jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    failed = False
    with cf.ThreadPoolExecutor(jobs) as ex:
        for src, obj, rc, err in ex.map(_compile, srcs):
            if err:
                print(f"[{os.path.basename(src)}]\n{err}", file=sys.stderr)
            if rc != 0:
                failed = True
            objs.append(obj)
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        print(p.stderr, file=sys.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


EXAMPLES = os.path.join(HERE, "examples")
CXX = os.environ.get("CXX", "g++")


def build_examples():
    """Host-only C++ drivers written against include/gpupfem2.hpp, linked with libpfem2_hip.so (found
    next to the binary's directory at run time)."""
    out = []
    for src in sorted(glob.glob(os.path.join(EXAMPLES, "*.cpp"))):
        exe = src[:-4]
        cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"), src,
               "-L" + HERE, "-lpfem2_hip", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib", "-o", exe + ".tmp"]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            print(p.stdout + p.stderr, file=sys.stderr)
            raise RuntimeError(f"example build failed: {os.path.basename(src)}")
        if p.stderr.strip():
            print(p.stderr, file=sys.stderr)
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


def build_oracle():
    p = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
    if p.returncode != 0:
        print(p.stdout + p.stderr, file=sys.stderr)
        raise RuntimeError("oracle build failed")
    return os.path.join(ROOT, "oracle", "liboracle.so")


if __name__ == "__main__":
    only_lib = "--lib" in sys.argv
    print(build_lib())
    print("\n".join(build_examples()))
    if not only_lib:
        print(build_oracle())
