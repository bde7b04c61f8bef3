"""Build libpartapp_ps.so (gfx950) in-tree with hipcc.

    python -m partapp_amd.build        # or partapp_amd.build.build()

The library lands in partapp_amd/lib/ next to this file, so a snapshot of the
repository carries it to a GPU machine without rebuilding.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIBNAME = "libpartapp_ps.so"

SOURCES = ["ps_api.hip", "ps_kernels.hip", "ps_offdiag.hip", "ps_localmax.hip", "ps_ingest.hip", "ps_mfma.hip", "ps_plan.cpp", "ps_synth.cpp", "ps_synth_dev.hip", "ps_matio.cpp", "ps_noalloc.hip"]

# -packed-fp32-ops: no v_pk_*_f32 from the compiler.  -ffp-contract=off: the
# only fused multiply-adds are the explicit fmaf() of the filters, and the
# host-side plan arithmetic is evaluated exactly as written.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-pthread",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + ".tmp"
    cmd = [hipcc()] + FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp]
    cmd += [os.path.join(CSRC, s) for s in SOURCES]
    cmd += ["-lz"]  # MAT-file compression (ps_matio.cpp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building %s" % LIBNAME)
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)
    os.replace(tmp, out)
    return out


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "wrapper_main.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "_build", "wrapper_main")


def build_cpp_tests(verbose=False):
    """The C++ driver of include/partapp_ps.hpp used by the GPU tests."""
    os.makedirs(os.path.dirname(CPP_TEST_BIN), exist_ok=True)
    tmp = CPP_TEST_BIN + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
           "-o", tmp, CPP_TEST_SRC, "-L", LIBDIR, "-lpartapp_ps", "-Wl,-rpath," + LIBDIR,
           "-Wl,-rpath,$ORIGIN/../../../partapp_amd/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building %s" % CPP_TEST_BIN)
    os.replace(tmp, CPP_TEST_BIN)
    return CPP_TEST_BIN


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_cpp_tests(verbose=True))
