"""GPU: the QLL C API (qll.h) driven from C the way QLL's testdslash.c and
testwilson.c drive it, checked against the FP64 oracle in QLL's host layout
(tests/qll_driver.c)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "qll_driver.c")
BIN = os.path.join(ROOT, "tests", "qll_driver.bin")


def _driver():
    if not os.path.exists(BIN) or os.path.getmtime(BIN) < os.path.getmtime(SRC):
        subprocess.run(["gcc", "-O1", "-fopenmp", "-std=gnu11", "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "oracle"), "-o", BIN, SRC,
                        "-L", os.path.join(ROOT, "qll_amd", "lib"), "-lqll_amd",
                        "-L", os.path.join(ROOT, "oracle"), "-loracle",
                        "-Wl,-rpath," + os.path.join(ROOT, "qll_amd", "lib"),
                        "-Wl,-rpath," + os.path.join(ROOT, "oracle"), "-lm"], check=True)
    return BIN


@pytest.mark.parametrize("prec,vlen,dims", [
    ("D", 1, (4, 4, 4, 8)),
    ("D", 4, (4, 4, 8, 8)),
    ("F", 8, (4, 4, 8, 8)),
    ("F", 1, (6, 4, 4, 8)),
])
def test_qll_api_driver(prec, vlen, dims):
    print("case: qll_driver %s vlen=%d dims=%s" % (prec, vlen, dims), flush=True)
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([_driver(), prec, str(vlen)] + [str(d) for d in dims], capture_output=True, text=True,
                       timeout=300, env=env)
    print(r.stdout[-6000:])
    print(r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-3000:]
    assert "PASSED" in r.stdout


TD_SRC = os.path.join(ROOT, "tests", "qll_testdslash.c")


def _testdslash(prec):
    """qll_testdslash.c built against libqll_amd and the oracle's qop* library
    of one precision (-DDOUBLE for D), as QLL builds its tests."""
    b = os.path.join(ROOT, "tests", "qll_testdslash_%s.bin" % prec)
    if not os.path.exists(b) or os.path.getmtime(b) < os.path.getmtime(TD_SRC):
        subprocess.run(["gcc", "-O1", "-fopenmp", "-std=gnu11"] + (["-DDOUBLE"] if prec == "D" else []) +
                       ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "oracle"), "-o", b, TD_SRC,
                        "-L", os.path.join(ROOT, "qll_amd", "lib"), "-lqll_amd",
                        "-L", os.path.join(ROOT, "oracle"), "-lqop" + prec, "-loracle",
                        "-Wl,-rpath," + os.path.join(ROOT, "qll_amd", "lib"),
                        "-Wl,-rpath," + os.path.join(ROOT, "oracle"), "-lm"], check=True)
    return b


@pytest.mark.parametrize("prec,vlen,dims", [
    ("D", 1, (4, 4, 4, 8)),
    ("D", 4, (4, 4, 8, 8)),
    ("F", 1, (4, 6, 4, 8)),
    ("F", 8, (4, 4, 8, 8)),
])
def test_qll_testdslash_vs_qop(prec, vlen, dims):
    """QLL's testdslash/testwilson flow: every API result against the qop*
    entry point the reference's tests compare with (here the FP64 oracle)."""
    print("case: qll_testdslash %s vlen=%d dims=%s" % (prec, vlen, dims), flush=True)
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([_testdslash(prec), str(vlen)] + [str(d) for d in dims], capture_output=True, text=True,
                       timeout=300, env=env)
    print(r.stdout[-8000:])
    print(r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-4000:]
    assert "PASSED" in r.stdout


FT_SRC = os.path.join(ROOT, "tests", "qll_fieldtest.c")


def _fieldtest(prec):
    """qll_fieldtest.c, a QLL program written against common.h and qll.h's
    generic names, built per precision like QLL's own drivers."""
    b = os.path.join(ROOT, "tests", "qll_fieldtest_%s.bin" % prec)
    if not os.path.exists(b) or os.path.getmtime(b) < os.path.getmtime(FT_SRC):
        subprocess.run(["gcc", "-O1", "-fopenmp", "-std=gnu11"] + (["-DDOUBLE"] if prec == "D" else []) +
                       ["-I", os.path.join(ROOT, "include"), "-o", b, FT_SRC,
                        "-L", os.path.join(ROOT, "qll_amd", "lib"), "-lqll_amd",
                        "-L", os.path.join(ROOT, "oracle"), "-lqop" + prec, "-loracle",
                        "-Wl,-rpath," + os.path.join(ROOT, "qll_amd", "lib"),
                        "-Wl,-rpath," + os.path.join(ROOT, "oracle"), "-lm"], check=True)
    return b


@pytest.mark.parametrize("prec,vlen,dims", [
    ("D", 1, (4, 4, 4, 8)),
    ("D", 4, (4, 4, 8, 8)),
    ("F", 8, (4, 4, 8, 8)),
    ("F", 1, (6, 4, 4, 8)),
])
def test_qll_fieldtest_transporters(prec, vlen, dims):
    """fieldtest.c:183-204 (transport forth and back, matrix transport vs
    qopTransportM), split-phase transporters in a team, makeShift for an
    arbitrary displacement, and a setOffSubC2/gsum Linop in cgSolve vs
    solve2."""
    print("case: qll_fieldtest %s vlen=%d dims=%s" % (prec, vlen, dims), flush=True)
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([_fieldtest(prec), str(vlen)] + [str(d) for d in dims], capture_output=True, text=True,
                       timeout=300, env=env)
    print(r.stdout[-6000:])
    print(r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-4000:]
    assert "PASSED" in r.stdout
