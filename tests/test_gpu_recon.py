"""GPU: asqtad links whose long (Naik) terms, and Wilson links whose terms,
are a scale times a unitary matrix (checked in FP64) are read as two rows and
the third is rebuilt in the Dslash, also in the FP32/FP16/Q16 copies converted
from them; other links are read in full.  Both give the oracle's Dslash to the precision's bound (FP64
1e-13, FP32 1e-5, FP16 5e-3, Q16 2e-4, relative L2 over the output parity)."""
import numpy as np
import pytest

import oracle as O
from helpers import parity_mask, rel_err

pytestmark = pytest.mark.gpu

TOL = {0: 1e-13, 1: 1e-5, 2: 5e-3, 3: 2e-4}


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
    assert L64.converted(gpu.F32).reconstructs                 # sloppy copies inherit the check
    assert L64.converted(gpu.Q16).reconstructs
    assert not L64.converted(gpu.Q16).converted(gpu.F64).reconstructs   # 16-bit round trip: not unitary to 1e-12
    assert L64.converted(gpu.F64).reconstructs
    W = gpu.Links(lat, gpu.WILSON, 0, 8).from_gauge(gpu.Gauge(lat).random_su3(2))
    assert W.reconstructs and W.converted(gpu.F16).reconstructs   # every Wilson term is SU(3)
    S8 = gpu.Links(lat, gpu.STAG, 0, 8).from_gauge(gpu.Gauge(lat).random_su3(3))
    assert not S8.reconstructs                                   # no long terms


def test_generated_links_reconstruct_at_full_size(gpu):
    """The 32^3x64 link set of the multishift benchmark (2M sites per parity,
    8.4M Gaussian matrices) passes the 1e-12 check everywhere: the generator's
    second Gram-Schmidt pass keeps even its worst-conditioned matrices unitary
    to a few eps, so the long-link triple products stay within the check."""
    lat = gpu.Lattice((32, 32, 32, 64))
    L64 = gpu.asqtad_links_device(lat, 2)
    assert L64.reconstructs


def _prec_outputs(gpu, lat, L, y, prec, nrhs=1):
    outs = []
    for p in (0, 1):
        yin = gpu.Field(lat, prec, 3, nrhs, 1 - p).set_lex(y)
        out = gpu.Field(lat, prec, 3, nrhs, p)
        gpu.dslash(L, out, yin)
        outs.append(out.get_lex())
    return outs


@pytest.mark.parametrize("prec", [1, 2, 3])
def test_sloppy_copies_rebuild_long_rows(gpu, prec):
    """FP32/FP16/Q16 copies of reconstructing FP64 links rebuild the third
    rows from their rounded first two; perturbed links converted the same way
    are read in full.  Both match the oracle to the precision's bound."""
    print("case: sloppy rebuild prec=%d" % prec, flush=True)
    dims = (6, 4, 8, 4)
    links = O.asqtad_links(dims, 23)
    bad = links.copy()
    bad[5, 7, 0:2] *= 1.001
    y = O.gauss_vector(dims, 24, 100, parity=-1)
    lat = gpu.Lattice(dims)
    for lk, rec in ((links, True), (bad, False)):
        L64 = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(lk)
        Ls = L64.converted(prec)
        assert L64.reconstructs == rec and Ls.reconstructs == rec
        outs = _prec_outputs(gpu, lat, Ls, y, prec)
        for p in (0, 1):
            ref = O.dslash(dims, lk, 0.0, y, sub="even" if p == 0 else "odd")
            err = rel_err(outs[p], ref, parity_mask(dims, p))
            assert err < TOL[prec], (rec, p, err)


def test_fp32_rebuilt_rows_agree_with_full_reads(gpu):
    """The FP32 copy converted from reconstructing FP64 links and FP32 links
    imported directly (read in full: no FP64 check) hold the same rounded
    first two rows; the rebuilt third row differs from the stored one by FP32
    rounding only."""
    dims = (8, 4, 4, 8)
    links = O.asqtad_links(dims, 25)
    y = O.gauss_vector(dims, 26, 100, parity=-1)
    lat = gpu.Lattice(dims)
    Lr = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links).converted(gpu.F32)
    Lf = gpu.Links(lat, gpu.STAG, gpu.F32, 16).set_lex(links)
    assert Lr.reconstructs and not Lf.reconstructs
    a, b = _prec_outputs(gpu, lat, Lr, y, 1), _prec_outputs(gpu, lat, Lf, y, 1)
    for p in (0, 1):
        assert rel_err(a[p], b[p], parity_mask(dims, p)) < 1e-6


@pytest.mark.parametrize("prec", [0, 1, 3])
def test_block_dslash_rebuilt_and_full(gpu, prec):
    """The block Dslash (4 rhs) gives the oracle's result with rebuilt long
    rows and with full reads."""
    print("case: block rebuild prec=%d" % prec, flush=True)
    dims = (4, 6, 4, 8)
    links = O.asqtad_links(dims, 27)
    bad = links.copy()
    bad[3, 9, 4:6] *= 0.999
    ys = np.stack([O.gauss_vector(dims, 70 + r, 100, parity=-1) for r in range(4)])
    lat = gpu.Lattice(dims)
    for lk, rec in ((links, True), (bad, False)):
        L = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(lk)
        if prec:
            L = L.converted(prec)
        assert L.reconstructs == rec
        outs = _prec_outputs(gpu, lat, L, ys, prec, nrhs=4)
        for p in (0, 1):
            for r in range(4):
                ref = O.dslash(dims, lk, 0.0, ys[r], sub="even" if p == 0 else "odd")
                err = rel_err(outs[p][r], ref, parity_mask(dims, p))
                assert err < TOL[prec], (rec, p, r, err)


@pytest.mark.parametrize("prec,nrhs", [(0, 1), (1, 1), (3, 1), (0, 8), (1, 16), (2, 8)])
def test_wilson_rebuilt_and_full(gpu, prec, nrhs):
    """Wilson links built from SU(3) fields rebuild every term's third row (in
    every launch shape, with and without the LDS link stage); perturbed links
    are read in full.  Both match the oracle; all shapes agree bit for bit."""
    print("case: wilson rebuild prec=%d nrhs=%d" % (prec, nrhs), flush=True)
    dims = (8, 8, 8, 8)
    links = O.wilson_doublestore(dims, O.random_su3_field(dims, 35))
    bad = links.copy()
    bad[3, 11, 0:2] *= 1.001
    ys = np.stack([O.gauss_vector(dims, 80 + r, 300, parity=-1, nreal=24) for r in range(nrhs)])
    lat = gpu.Lattice(dims)
    m = parity_mask(dims, 0)
    for lk, rec in ((links, True), (bad, False)):
        L = gpu.Links(lat, gpu.WILSON, 0, 8).set_lex(lk)
        if prec:
            L = L.converted(prec)
        assert L.reconstructs == rec
        yin = gpu.Field(lat, prec, 12, nrhs, 1).set_lex(ys)
        res = []
        for mode in [0] + list(range(2, 20)) + [1]:
            gpu.set_tuning(mode)
            out = gpu.Field(lat, prec, 12, nrhs, 0)
            gpu.dslash(L, out, yin, b=-1.0)
            res.append(out.get_lex())
        gpu.set_tuning(1)
        for o in res[1:]:
            assert np.array_equal(o, res[0])
        got = res[0] if nrhs > 1 else res[0][None]
        for r in range(nrhs):
            ref = O.wilson_dslash(dims, lk, 0.0, ys[r], sub="even")
            assert rel_err(got[r], ref, m) < TOL[prec], (rec, r)
