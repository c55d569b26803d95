"""GPU: argument errors are reported through the C ABI (return code -1 and
qamd_last_error, QllAmdError in Python) and leave the library usable; limits
at their maximum are accepted."""
import numpy as np
import pytest

import oracle as O
from helpers import parity_mask, rel_err

pytestmark = pytest.mark.gpu


def _raises(gpu, fn, text):
    with pytest.raises(gpu.QllAmdError) as e:
        fn()
    assert text in str(e.value), str(e.value)


def test_object_creation_errors(gpu):
    lat = gpu.Lattice((4, 4, 4, 4))
    _raises(gpu, lambda: gpu.Lattice((4, 4, 4, 5)), "extents must be even")
    _raises(gpu, lambda: gpu.Lattice((2, 4, 4, 4)), "extents must be even and >= 4")
    _raises(gpu, lambda: gpu.Field(lat, 7, 3, 1, 0), "bad precision")
    _raises(gpu, lambda: gpu.Field(lat, 0, 6, 1, 0), "ncomp must be 3 or 12")
    _raises(gpu, lambda: gpu.Field(lat, 0, 3, 0, 0), "nrhs must be 1..64")
    _raises(gpu, lambda: gpu.Field(lat, 0, 3, 65, 0), "nrhs must be 1..64")
    _raises(gpu, lambda: gpu.Field(lat, 0, 3, 1, 2), "parity must be 0 or 1")
    _raises(gpu, lambda: gpu.Links(lat, gpu.WILSON, 0, 16), "Wilson links need 8 terms")
    _raises(gpu, lambda: gpu.Links(lat, gpu.STAG, 0, 12), "8 or 16 terms")
    f = gpu.Field(lat, 0, 3, 64, 1)   # the largest block is accepted
    assert f.nrhs == 64


def test_dslash_argument_errors(gpu):
    dims = (4, 4, 4, 8)
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(O.asqtad_links(dims, 3))
    W = gpu.Links(lat, gpu.WILSON, 0, 8).set_lex(O.wilson_doublestore(dims, O.random_su3_field(dims, 4)))
    e3, o3 = gpu.Field(lat, 0, 3, 1, 0), gpu.Field(lat, 0, 3, 1, 1)
    e12, o12 = gpu.Field(lat, 0, 12, 1, 0), gpu.Field(lat, 0, 12, 1, 1)
    _raises(gpu, lambda: gpu.dslash(L, e3, e3), "opposite parity")
    _raises(gpu, lambda: gpu.dslash(L, e3, o3, w=o3), "w must have the output parity")
    _raises(gpu, lambda: gpu.dslash(L, e3, gpu.Field(lat, 1, 3, 1, 1)), "do not fit links")
    _raises(gpu, lambda: gpu.dslash(W, e3, o3), "spinor fields")
    _raises(gpu, lambda: gpu.dslash(W, e12, e12), "parity mismatch")
    _raises(gpu, lambda: gpu.dslash(W, e12, gpu.Field(lat, 1, 12, 1, 1)), "precision does not match")
    _raises(gpu, lambda: gpu.dslash(W, e12, gpu.Field(lat, 0, 12, 2, 1)), "equal rhs counts")
    L32 = L.converted(1)
    _raises(gpu, lambda: gpu.dslash(L32, gpu.Field(lat, 1, 3, 17, 0), gpu.Field(lat, 1, 3, 17, 1)),
            "field shapes do not match")
    _raises(gpu, lambda: gpu.stag_cg(W, e3, e3, 0.1, 1e-10, 10), "staggered links required")
    LQ = L.converted(gpu.Q16)
    eq, bq = gpu.Field(lat, gpu.Q16, 3, 1, 0), gpu.Field(lat, gpu.Q16, 3, 1, 0)
    _raises(gpu, lambda: gpu.stag_cg(LQ, eq, bq, 0.1, 1e-10, 10), "must be FP64 or FP32")
    _raises(gpu, lambda: gpu.stag_cg(L.converted(gpu.F16), gpu.Field(lat, gpu.F16, 3, 1, 0),
                                     gpu.Field(lat, gpu.F16, 3, 1, 0), 0.1, 1e-10, 10), "must be FP64 or FP32")
    _raises(gpu, lambda: gpu.stag_cgms(L, [e3] * 33, e3, 0.1, [0.0] * 33, 1e-10, 10), "shifts")
    _raises(gpu, lambda: gpu.stag_cg_mixed_block(L, L.converted(1), gpu.Field(lat, 0, 3, 9, 0),
                                                 gpu.Field(lat, 0, 3, 9, 0), 0.1, 1e-10, 10), "1..8 rhs")
    _raises(gpu, lambda: gpu.wilson_cgne(L, e12, e12, 0.1, 1e-10, 10), "Wilson links required")
    _raises(gpu, lambda: gpu.stag_cg_mixed_batch([], 1e-10, 10), "problems")
    # the library stays usable after the errors: a valid call still matches the oracle
    y = O.gauss_vector(dims, 5, 100, parity=1)
    o3.set_lex(y)
    gpu.dslash(L, e3, o3)
    ref = O.dslash(dims, O.asqtad_links(dims, 3), 0.0, y, sub="even")
    assert rel_err(e3.get_lex(), ref, parity_mask(dims, 0)) < 1e-13


def test_largest_rhs_blocks(gpu):
    """The largest FP32 block (16 rhs) in one block-Dslash launch and a
    16-rhs Wilson block give every rhs the oracle's result."""
    dims = (4, 4, 4, 8)
    links = O.asqtad_links(dims, 6)
    lat = gpu.Lattice(dims)
    L32 = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links).converted(1)
    ys = np.stack([O.gauss_vector(dims, 80 + r, 100, parity=1) for r in range(16)])
    yin = gpu.Field(lat, 1, 3, 16, 1).set_lex(ys)
    out = gpu.Field(lat, 1, 3, 16, 0)
    gpu.dslash(L32, out, yin)
    got = out.get_lex()
    m = parity_mask(dims, 0)
    for r in range(16):
        assert rel_err(got[r], O.dslash(dims, links, 0.0, ys[r], sub="even"), m) < 1e-5, r
