"""GPU: qamd_dslash_noalloc, the allocation-free form of qamd_dslash (partial
sums of |out|^2 in a fixed device array instead of library scratch), called
through the C ABI and compared bit for bit with qamd_dslash: two lattices in
a row, the second smaller, so the second call reads none of the first
call's leftovers; both against the oracle's Dslash."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from helpers import parity_mask, rel_err

pytestmark = pytest.mark.gpu

TOL = {0: 1e-13, 1: 1e-5, 3: 2e-4}


def _entry(gpu):
    f = gpu.lib().qamd_dslash_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_double,
                  C.POINTER(C.c_double)]
    return f


@pytest.mark.parametrize("prec", [0, 1, 3])
def test_noalloc_dslash_matches_dslash(gpu, prec):
    f = _entry(gpu)
    for dims, seed in (((8, 8, 8, 16), 3), ((4, 4, 6, 8), 4)):
        print("case: dslash_noalloc prec=%d dims=%s" % (prec, dims), flush=True)
        links = O.asqtad_links(dims, seed)
        y = O.gauss_vector(dims, seed + 1, 100, parity=-1)
        lat = gpu.Lattice(dims)
        L = gpu.Links(lat, gpu.STAG, gpu.F64, 16).set_lex(links)
        if prec:
            L = L.converted(prec)
        for p in (0, 1):
            yin = gpu.Field(lat, prec, 3, 1, 1 - p).set_lex(y)
            w = gpu.Field(lat, prec, 3, 1, p).set_lex(y)
            ref = gpu.Field(lat, prec, 3, 1, p)
            n_ref = gpu.dslash(L, ref, yin, a=0.3, w=w, b=-1.0, want_norm=True)
            out = gpu.Field(lat, prec, 3, 1, p)
            n = C.c_double(-1.0)
            assert f(L.h, out.h, yin.h, 0.3, w.h, -1.0, C.byref(n)) == 0, gpu.lib().qamd_last_error()
            got = out.get_lex()
            assert np.array_equal(got, ref.get_lex())
            assert n.value == n_ref
            want = 0.3 * y - O.dslash(dims, links, 0.0, y, sub="even" if p == 0 else "odd")
            assert rel_err(got, want, parity_mask(dims, p)) < TOL[prec]
            # without the reduction: same output, no scratch touched
            out2 = gpu.Field(lat, prec, 3, 1, p)
            assert f(L.h, out2.h, yin.h, 0.3, w.h, -1.0, None) == 0
            assert np.array_equal(out2.get_lex(), got)


def test_noalloc_dslash_rejects_what_it_does_not_cover(gpu):
    f = _entry(gpu)
    dims = (4, 4, 4, 8)
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.STAG, gpu.F64, 16).set_lex(O.asqtad_links(dims, 5))
    W = gpu.Links(lat, gpu.WILSON, gpu.F64, 8).set_lex(O.wilson_doublestore(dims, O.random_su3_field(dims, 6)))
    n = C.c_double()
    blk_in, blk_out = gpu.Field(lat, 0, 3, 2, 1), gpu.Field(lat, 0, 3, 2, 0)
    assert f(L.h, blk_out.h, blk_in.h, 0.0, None, 1.0, C.byref(n)) != 0
    assert b"single-rhs" in gpu.lib().qamd_last_error()
    s_in, s_out = gpu.Field(lat, 0, 12, 1, 1), gpu.Field(lat, 0, 12, 1, 0)
    assert f(W.h, s_out.h, s_in.h, 0.0, None, 1.0, C.byref(n)) != 0
    assert b"staggered" in gpu.lib().qamd_last_error()
    assert f(None, s_out.h, s_in.h, 0.0, None, 1.0, C.byref(n)) != 0
