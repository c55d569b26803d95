"""The C-ABI library loads and exports every function include/*.h declares
(CPU only: no compute call is made)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "qll_amd", "lib", "libqll_amd.so")

# declared by QLL but never defined there either (include/qllP.h:71-74,108-109)
NOT_IN_REFERENCE = {"solveF", "solveD", "solveMultiF", "solveMultiD", "wilsonSolveMultiF", "wilsonSolveMultiD"}


def declared_symbols():
    names = set()
    # per-precision prototypes: PR(name)(...)
    for hdr in ("qllP.h", "solversP.h"):
        txt = open(os.path.join(ROOT, "include", hdr)).read()
        for m in re.finditer(r"PR\((\w+)\)\s*\(", txt):
            if m.group(1) not in ("Linop",):
                names.add(m.group(1) + "F")
                names.add(m.group(1) + "D")
    for hdr in ("qll.h", "layout.h", "thread.h", "qll_amd.h", "common.h"):
        txt = open(os.path.join(ROOT, "include", hdr)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"^\s*#(?:[^\n]*\\\n)*[^\n]*", "", txt, flags=re.M)   # directives, with continuations
        for m in re.finditer(r"^\s*[\w\s\*]+?\b(\w+)\s*\([^;{]*\)\s*;", txt, flags=re.M):
            n = m.group(1)
            if n not in ("PR", "P", "defined") and not n.startswith("qop"):
                names.add(n)
    return names - NOT_IN_REFERENCE


# The reference's exported API, listed from its headers and definitions (not
# from this repository's headers, so a symbol missing here fails the test):
# include/qllP.h:8-109, include/solversP.h:16-22, include/common.h:51-64 and
# 88-91, lib/cg.c:37 (cglsSet), lib/common.c:10,53, lib/layout.c:6-217, and
# the vla.h functions behind the generic names of include/qll.h:56-57 and
# include/vla.h:120-140 (X_veq_r, norm2_X, ...) with makeU (vla.h:85).
REFERENCE_PER_PREC = """
fieldNew fieldFree fieldSet fieldZero fieldSumElem fieldNorm2
transporterNew transporterFree transporterStartT transporterLocalT transporterWaitT transporterDoT transporterDo
gaugeDoublestore gaugePlaq smearFat7
stagDslashSetup createStaggeredLinks createStaggeredLinksFb freeStaggeredLinks dslash2
createStaggeredSolver freeStaggeredSolver solve2 solveMulti2
createWilsonLinks freeWilsonLinks wilsonDslash createWilsonSolver freeWilsonSolver wilsonSolve
cglsInit cglsFree cglsSet cgSolve cgmsSolve
shift getShift getStaggeredSI dslashResetTimer dslashPrintTimer wilsonDslashResetTimer wilsonDslashPrintTimer
X_veq_r X_veq_r_times_X X_vpeq_r_times_X X_veq_X_plus_r_times_X sumElem_X norm2_X r_veq_norm2_X
re_X_dot_X r_veq_re_X_dot_X makeU
""".split()
REFERENCE_UNSUFFIXED = """
myalloc gsum myrank layoutSetup layoutIndex layoutCoord layoutShift layoutSubset
makeShift makeShiftSub makeShiftMulti makeShiftMultiSub
""".split()


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l or " D " in l or " B " in l}


def test_reference_api_exported():
    want = {n + s for n in REFERENCE_PER_PREC for s in "FD"} | set(REFERENCE_UNSUFFIXED)
    missing = sorted(want - _exported())
    assert not missing, missing


def test_library_built_and_loads():
    assert os.path.exists(LIB), "run __graft_entry__.build() first"
    ctypes.CDLL(LIB)


def test_every_declared_symbol_exported():
    exported = _exported()
    missing = sorted(declared_symbols() - exported)
    assert not missing, missing
    assert "dslash2D" in exported and "qamd_dslash" in exported


def test_no_device_fails_loudly():
    """Without a usable GPU every entry point reports an error; nothing falls
    back to a host computation."""
    import qll_amd
    try:
        qll_amd.init()
    except qll_amd.QllAmdError as e:
        assert "device" in str(e)
    else:
        pytest.skip("a GPU is present")


def test_lattice_arguments_checked_before_any_device_call():
    """Odd or too small extents are refused with a message (no GPU needed)."""
    lib = ctypes.CDLL(LIB)
    lib.qamd_lattice_create.restype = ctypes.c_void_p
    lib.qamd_last_error.restype = ctypes.c_char_p
    for dims in ((4, 4, 4, 5), (2, 4, 4, 4), (8, 8, 0, 8)):
        assert not lib.qamd_lattice_create((ctypes.c_int * 4)(*dims))
        assert b"extents must be even and >= 4" in lib.qamd_last_error()
    assert not lib.qamd_lattice_create((ctypes.c_int * 4)(2048, 2048, 1024, 1024))
    assert b"site limit" in lib.qamd_last_error()
