"""GPU: FP64 asqtad links whose long (Naik) terms are a scale times a unitary
matrix are read as two rows and the third is rebuilt in the Dslash; other
links are read in full.  Both give the oracle's Dslash to 1e-13."""
import numpy as np
import pytest

import oracle as O
from helpers import parity_mask, rel_err

pytestmark = pytest.mark.gpu


def _dslash_err(gpu, dims, links, seed):
    y = O.gauss_vector(dims, seed, 100, parity=-1)
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    errs = []
    for p in (0, 1):
        yin = gpu.Field(lat, 0, 3, 1, 1 - p).set_lex(y)
        out = gpu.Field(lat, 0, 3, 1, p)
        gpu.dslash(L, out, yin)
        ref = O.dslash(dims, links, 0.0, y, sub="even" if p == 0 else "odd")
        errs.append(rel_err(out.get_lex(), ref, parity_mask(dims, p)))
    return L, max(errs)


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (6, 4, 8, 4), (8, 8, 8, 8)])
def test_rebuilt_long_rows_match_oracle(gpu, dims):
    links = O.asqtad_links(dims, 17)      # fat7 + Naik from SU(3) with phases: both det signs occur
    L, err = _dslash_err(gpu, dims, links, 18)
    assert L.reconstructs
    assert err < 1e-13, err


def test_general_long_links_read_in_full(gpu):
    dims = (4, 4, 6, 4)
    links = O.asqtad_links(dims, 19).copy()
    links[5, 7, 0:2] *= 1.001              # one entry of one long term at one site: no longer a scaled unitary
    L, err = _dslash_err(gpu, dims, links, 20)
    assert not L.reconstructs
    assert err < 1e-13, err
    rnd = np.random.default_rng(3).standard_normal(links.shape)   # not unitary anywhere
    L2, err2 = _dslash_err(gpu, dims, rnd, 21)
    assert not L2.reconstructs and err2 < 1e-13


def test_device_links_and_conversions(gpu):
    lat = gpu.Lattice((8, 8, 8, 8))
    L64 = gpu.asqtad_links_device(lat, 4)
    assert L64.reconstructs
    assert not L64.converted(gpu.F32).reconstructs             # FP32 links are read in full
    assert not L64.converted(gpu.Q16).converted(gpu.F64).reconstructs   # 16-bit round trip: not unitary to 1e-12
    assert L64.converted(gpu.F64).reconstructs
    W = gpu.Links(lat, gpu.WILSON, 0, 8).from_gauge(gpu.Gauge(lat).random_su3(2))
    assert not W.reconstructs
