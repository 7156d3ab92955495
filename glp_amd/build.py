"""Builds libglp.so (HIP kernels for gfx950 + C++ host layer) and the
command-line programs, in-tree, with hipcc and g++ directly.

    python -m glp_amd.build          # or glp_amd.build.build()
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("GLP_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["kernels.hip", "counting.hip", "session.hip", "batch.hip", "server.hip", "plan.hip",
               "autotune.hip", "gen_device.hip", "noalloc.hip"]
CPP_SOURCES = ["graph_io.cpp", "gen.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = COMMON + [
    "--offload-arch=" + ARCH,
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]
CPP_FLAGS = COMMON + ["-fopenmp", "-I" + os.path.join(ROOT, "include")]

LIB = os.path.join(HERE, "libglp.so")
BIN = os.path.join(HERE, "bin")
PROGRAMS = {
    "GPULP": ["gpulp_main.cpp"],
    "edgelist2bin": ["tool_edgelist2bin.cpp"],
    "normalize": ["tool_normalize.cpp"],
    "direct_edgelist2bin": ["tool_direct_edgelist2bin.cpp"],
    "glp_autotune": ["autotune_main.cpp"],
}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), r.stdout))
    return r.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", out]
    else:
        cmd = ["g++"] + CPP_FLAGS + ["-c", path, "-o", out]
    _run(cmd)
    return out


def build(verbose=False):
    """Compiles every source (always) and links libglp.so and the programs."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    srcs = [s for s in HIP_SOURCES + CPP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = int(os.environ.get("MAX_JOBS", "8"))
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, 16))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run(["g++", "-shared", "-o", tmp] + objs +
         ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-fopenmp",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    for prog, files in PROGRAMS.items():
        paths = [os.path.join(CSRC, "host", f) for f in files]
        if not all(os.path.exists(p) for p in paths):
            continue
        _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
              "-I" + os.path.join(CSRC, "host")] + paths +
             ["-o", os.path.join(BIN, prog), "-L" + HERE, "-lglp", "-Wl,-rpath,$ORIGIN/..",
              "-fopenmp"])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
