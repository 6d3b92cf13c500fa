"""Builds libcubble_amd.so (HIP, gfx950) and the cubble-cli executable in-tree.

    python -m cubble_amd.build          # from the repository root

Every HIP source is compiled with -ffp-contract=off except the mixed-precision
pair kernel, so that the FP64 path rounds exactly as written; packed FP32 ops
are disabled for all device code.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcubble_amd.so")
CLI = os.path.join(HERE, "cubble-cli")

ARCH = os.environ.get("CUBBLE_AMD_ARCH", "gfx950")

HIP_SOURCES = ["step_kernels.hip", "pair_mixed.hip", "search_kernels.hip", "remove_kernels.hip",
               "engine.hip", "state_io.hip", "runtime.hip", "server.hip", "autotune.hip",
               "observe_kernels.hip", "pairs.hip", "pairs_noalloc.hip"]
HOST_SOURCES = ["host_init.cpp", "cli.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-I", INCLUDE, "-I", CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src, verbose):
    hipcc = _hipcc()
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src.endswith(".hip"):
        contract = "-ffp-contract=fast" if src == "pair_mixed.hip" else "-ffp-contract=off"
        cmd = [hipcc, "-c", path, "-o", out] + COMMON + DEVICE + [contract]
    else:
        cmd = [hipcc, "-c", path, "-o", out] + COMMON + ["-ffp-contract=off"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (src, r.stdout, r.stderr))
    return out


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    srcs = HIP_SOURCES + HOST_SOURCES
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc, "-shared", "-fPIC", "-o", tmp] + objs + ["--offload-arch=" + ARCH]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    main_src = os.path.join(CSRC, "cli_main.cpp")
    cmd = [hipcc, main_src, "-o", CLI, "-I", INCLUDE, "-L", HERE, "-lcubble_amd",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("cli link failed:\n%s\n%s" % (r.stdout, r.stderr))
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
