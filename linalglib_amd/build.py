"""Builds the gfx950 library liblal.so in-tree with hipcc.

    python -m linalglib_amd.build          # compile every csrc/*.hip and link

The library holds the HIP kernels, the C ABI declared in include/lal.h and
the C++ host code above it (both in linalglib_amd/csrc/*.hip; plain *.cpp
sources there, if any, are compiled as host code).  Flags:

* ``--offload-arch=gfx950`` only (no other targets);
* ``-ffp-contract=off`` so that no multiply-add is fused unless the source
  says ``fma`` (the setup kernels must round exactly like the host path);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops`` keeps the compiler
  from emitting packed FP32 instructions.
"""
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIB_DIR, "liblal.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip", *DEVICE, *COMMON, "-c", src, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, (r.stdout + r.stderr)


def build(verbose=False, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    objs = []
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, out in ex.map(_compile, srcs):
            msgs = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln
                    and "is not a recognized feature" not in ln]
            if rc != 0:
                failed.append((src, "\n".join(msgs)))
            elif verbose and msgs:
                print("\n".join(msgs))
            objs.append(obj)
    if failed:
        for src, out in failed:
            print(f"--- {src}\n{out}", file=sys.stderr)
        raise RuntimeError(f"hipcc failed on {len(failed)} file(s)")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stdout + r.stderr, file=sys.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


def build_cpp_api_test():
    """The host C++ test of include/lal_linalg.hpp (g++ only, links liblal.so)."""
    src = os.path.join(ROOT, "tests", "cpp", "test_linalg_api.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "test_linalg_api")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}", src,
           f"-L{LIB_DIR}", "-llal", "-Wl,-rpath,$ORIGIN/../../linalglib_amd/_lib", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stdout + r.stderr, file=sys.stderr)
        raise RuntimeError("g++ failed on the C++ API test")
    return out


def build_mtx_tool():
    """Host-only Matrix Market driver for the CPU parity tests (g++ only)."""
    src = os.path.join(ROOT, "tests", "cpp", "mtx_tool.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "mtx_tool")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}", src,
           f"-L{LIB_DIR}", "-llal", "-Wl,-rpath,$ORIGIN/../../linalglib_amd/_lib", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stdout + r.stderr, file=sys.stderr)
        raise RuntimeError("g++ failed on the Matrix Market tool")
    return out


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_cpp_api_test())
    print(build_mtx_tool())
