"""Build libdda_hip.so for gfx950 with hipcc (in-tree, no JIT cache).

Each HIP source is compiled to an object in dda_amd/_build/ in parallel, then
linked into dda_amd/_lib/libdda_hip.so.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "_lib", "libdda_hip.so")
SOURCES = ["api.hip", "matvec.hip", "krylov.hip", "observe.hip", "server.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off (correctness on this hardware pool)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cpp"):  # host-only C++
        cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-pthread", "-c", os.path.join(SRC, src), "-o", obj]
    else:
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, out))
    return obj, out


def build_adapter_check():
    """The C++ adapter driver used by the GPU tests (g++ against the header)."""
    root = os.path.dirname(HERE)
    out = os.path.join(BUILD, "adapter_check")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-o", out, os.path.join(root, "tests", "cpp", "adapter_check.cpp"),
           "-L" + os.path.dirname(LIB), "-ldda_hip", "-Wl,-rpath,$ORIGIN/../_lib"]
    subprocess.check_call(cmd)
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(_compile, SOURCES))
    for _, out in results:
        if out.strip() and verbose:
            print(out)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + [o for o, _ in results]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_adapter_check()
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
