"""Build recipe for libqll_amd.so (gfx950 HIP kernels + C host API).

    python -m qll_amd.build          # or qll_amd.build.build()

HIP sources (csrc/*.hip) are compiled with hipcc for gfx950 only; the QLL
C API (csrc/host/*.c) with gcc as C99/C11; everything is linked into
qll_amd/lib/libqll_amd.so, which travels with the source tree.
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
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-fopenmp", "-lm"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
