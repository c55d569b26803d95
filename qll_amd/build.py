"""Build recipe for libqll_amd.so (gfx950 HIP kernels + C host API).

    python -m qll_amd.build          # or qll_amd.build.build()

HIP sources (csrc/*.hip) are compiled with hipcc for gfx950 only; the QLL
C API (csrc/host/*.c) with gcc as C99/C11; everything is linked into
qll_amd/lib/libqll_amd.so, which travels with the source tree.  The QLL-API
timing driver (csrc/bench/qll_api_bench.c) is built beside it.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libqll_amd.so")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIPFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-pass-failed",
]
CFLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function", "-fopenmp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + os.path.basename(cmd[-1]))
    return r.stderr


def build(verbose=False, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    jobs_list = []
    for src in sorted(glob.glob(os.path.join(CSRC, "*.hip"))):
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        jobs_list.append([HIPCC] + HIPFLAGS + inc + ["-c", "-o", obj, src])
    for src in sorted(glob.glob(os.path.join(CSRC, "host", "*.c"))):
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        jobs_list.append(["gcc"] + CFLAGS + inc + ["-c", "-o", obj, src])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd, err in zip(jobs_list, ex.map(_run, jobs_list)):
            if verbose and err.strip():
                sys.stderr.write(err)
    objs = [c[c.index("-o") + 1] for c in jobs_list]
    tmp = LIB + ".tmp"
    # The OpenMP teams of the C host API (BEGIN_PARALLEL, gsum, and the
    # caller's Linop callbacks that run inside them) must use the same
    # runtime as the QLL programs calling it, which gcc -fopenmp builds
    # against GNU libgomp; hipcc's -fopenmp would link LLVM's libomp, whose
    # teams do not see the caller's libgomp barriers.
    gomp = subprocess.run(["gcc", "-print-file-name=libgomp.so"], capture_output=True, text=True,
                          check=True).stdout.strip()
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + [gomp, "-lm"])
    os.replace(tmp, LIB)
    build_api_bench()
    return LIB


API_BENCH_SRC = os.path.join(CSRC, "bench", "qll_api_bench.c")
API_BENCH = os.path.join(LIBDIR, "qll_api_bench.bin")


def build_api_bench():
    """The QLL-API timing driver bench.py runs (a C program against the QLL
    headers and libqll_amd.so, found next to it through its rpath)."""
    _run(["gcc"] + CFLAGS + ["-I", os.path.join(ROOT, "include"), "-o", API_BENCH, API_BENCH_SRC,
                             "-L", LIBDIR, "-lqll_amd", "-Wl,-rpath,$ORIGIN", "-lm"])
    return API_BENCH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
