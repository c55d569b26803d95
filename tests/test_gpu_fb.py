"""GPU parity: the forward-only asqtad Dslash (link kind STAG_FB, QLL's
createStaggeredLinksFb / dslashFb2T, lib/stagDslash.c:110-176, 464-519)
against the oracle's restatement of dslashFb2T.

Tolerances (relative L2 difference over the output parity) as for the
doublestored Dslash: FP64 1e-13, FP32 1e-5, FP16 5e-3, Q16 2e-4.
"""
import numpy as np
import pytest

import oracle as O
from helpers import parity_mask, rel_err

pytestmark = pytest.mark.gpu

TOL = {0: 1e-13, 1: 1e-5, 2: 5e-3, 3: 2e-4}
DIMS = [(4, 4, 4, 4), (8, 8, 8, 8), (4, 6, 8, 10), (6, 4, 4, 12)]


def _fb_case(dims, seed, naik=True):
    u = O.apply_stag_phases(dims, O.random_su3_field(dims, seed))
    fl, ll = O.fat7(dims, u, u, want_long=naik)
    y = O.gauss_vector(dims, seed + 1, 100, parity=-1)
    return fl, ll, y


def _fb_links(gpu, lat, fl, ll, prec=0):
    terms = np.concatenate([fl, ll]) if ll is not None else fl
    L = gpu.Links(lat, gpu.STAG_FB, 0, terms.shape[0]).set_lex(terms)
    return L if prec == 0 else L.converted(prec)


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_fb_eo_dslash(gpu, dims, prec):
    print("case: forward-only asqtad EO dims=%s prec=%d" % (dims, prec), flush=True)
    fl, ll, y = _fb_case(dims, 21)
    lat = gpu.Lattice(dims)
    L = _fb_links(gpu, lat, fl, ll, prec)
    assert L.reconstructs   # fat7 long links of SU(3) fields are 3 x unitary
    for p in (0, 1):
        yin = gpu.Field(lat, prec, 3, 1, 1 - p).set_lex(y)
        out = gpu.Field(lat, prec, 3, 1, p)
        gpu.dslash(L, out, yin)
        ref = O.dslash_fb(dims, fl, ll, 0.0, y, sub="even" if p == 0 else "odd")
        err = rel_err(out.get_lex(), ref, parity_mask(dims, p))
        assert err < TOL[prec], (p, err)


@pytest.mark.parametrize("naik", [False, True])
def test_fb_dslash_all_with_mass(gpu, naik):
    """dslash2 on 'all' with Fb links: x = mass*y + D y, both parities, and
    the |out|^2 reduction."""
    dims = (4, 4, 6, 8)
    fl, ll, y = _fb_case(dims, 5, naik)
    mass = 0.37
    lat = gpu.Lattice(dims)
    L = _fb_links(gpu, lat, fl, ll)
    ref = O.dslash_fb(dims, fl, ll, mass, y, sub="all")
    got = np.zeros_like(y)
    for p in (0, 1):
        yo = gpu.Field(lat, 0, 3, 1, 1 - p).set_lex(y)
        yp = gpu.Field(lat, 0, 3, 1, p).set_lex(y)
        out = gpu.Field(lat, 0, 3, 1, p)
        n2 = gpu.dslash(L, out, yo, a=mass, w=yp, b=1.0, want_norm=True)
        got = out.get_lex(got)
        m = parity_mask(dims, p)
        assert abs(n2 - np.sum(ref[m] ** 2)) <= 1e-12 * np.sum(ref[m] ** 2)
    assert rel_err(got, ref) < 1e-13


def test_fb_full_rows_without_rebuild(gpu):
    """Long links that are not a scale times a unitary matrix are read in
    full (no third-row rebuild) and still match."""
    dims = (4, 4, 4, 8)
    fl, ll, y = _fb_case(dims, 8)
    rng = np.random.default_rng(3)
    ll = ll + 0.1 * rng.standard_normal(ll.shape)
    lat = gpu.Lattice(dims)
    L = _fb_links(gpu, lat, fl, ll)
    assert not L.reconstructs
    for p in (0, 1):
        yin = gpu.Field(lat, 0, 3, 1, 1 - p).set_lex(y)
        out = gpu.Field(lat, 0, 3, 1, p)
        gpu.dslash(L, out, yin)
        ref = O.dslash_fb(dims, fl, ll, 0.0, y, sub="even" if p == 0 else "odd")
        assert rel_err(out.get_lex(), ref, parity_mask(dims, p)) < 1e-13


@pytest.mark.parametrize("prec", [0, 1, 3])
def test_fb_device_links_half_storage(gpu, prec):
    """Forward-only links built on the GPU (fat7 -> from_gauge) give the
    doublestored set's Dslash with half its link bytes."""
    dims = (8, 8, 8, 16)
    lat = gpu.Lattice(dims)
    Lds = gpu.asqtad_links_device(lat, 77, prec=prec)
    Lfb = gpu.asqtad_links_device(lat, 77, prec=prec, kind=gpu.STAG_FB)
    assert Lds.reconstructs and Lfb.reconstructs
    V = O.volume(dims)
    sign_bytes = V   # one determinant-sign byte per site (lattice padded to 64 sites)
    assert Lfb.nbytes - sign_bytes == (Lds.nbytes - sign_bytes) // 2, (Lfb.nbytes, Lds.nbytes)
    y = O.gauss_vector(dims, 4, 100, parity=-1)
    for p in (0, 1):
        yin = gpu.Field(lat, prec, 3, 1, 1 - p).set_lex(y)
        a = gpu.Field(lat, prec, 3, 1, p)
        b = gpu.Field(lat, prec, 3, 1, p)
        gpu.dslash(Lds, a, yin)
        gpu.dslash(Lfb, b, yin)
        m = parity_mask(dims, p)
        # same links, a different summation order: agree to the storage precision
        assert rel_err(b.get_lex(), a.get_lex(), m) < {0: 1e-13, 1: 1e-6, 3: 1e-4}[prec]


def _true_rel_resid(dims, links, mass, b, x):
    Ax, _ = O.normal_op(dims, links, mass, x)
    m = parity_mask(dims, 0)
    return np.sum((b - Ax)[m] ** 2) / np.sum(b[m] ** 2)


def test_fb_cg_matches_oracle(gpu):
    """cgSolveT (lib/cg.c:43-122) on Fb links: the oracle CG's iteration count
    and solution on the equivalent doublestored set."""
    dims = (8, 8, 8, 8)
    fl, ll, _ = _fb_case(dims, 31)
    links = O.stag_doublestore(dims, fl, ll)
    b = O.gauss_vector(dims, 2, 50, parity=0)
    mass, rsq = 0.2, 1e-16
    xr, its_ref, _ = O.cg(dims, links, mass, b, rsq, 2000, noscale=True)
    lat = gpu.Lattice(dims)
    L = _fb_links(gpu, lat, fl, ll)
    x = gpu.Field(lat, 0, 3, 1, 0)
    bf = gpu.Field(lat, 0, 3, 1, 0).set_lex(b)
    info = gpu.stag_cg(L, x, bf, mass, rsq, 2000)
    assert info["converged"]
    assert abs(info["iterations"] - its_ref) <= 1, (info["iterations"], its_ref)
    xs = x.get_lex()
    assert _true_rel_resid(dims, links, mass, b, xs) < 4 * rsq
    assert rel_err(xs, xr, parity_mask(dims, 0)) < 1e-6


@pytest.mark.parametrize("sloppy", [1, 3])
def test_fb_mixed_cg(gpu, sloppy):
    """Mixed-precision CG on forward-only FP64 and sloppy (FP32 / Q16) links
    reaches the FP64 true residual 1e-10 (the oracle's operator) and the
    oracle's solution."""
    dims = (8, 8, 8, 8)
    mass = 0.05
    fl, ll, _ = _fb_case(dims, 12)
    links = O.stag_doublestore(dims, fl, ll)
    b = O.gauss_vector(dims, 13, 100, parity=0)
    lat = gpu.Lattice(dims)
    L64 = _fb_links(gpu, lat, fl, ll)
    Ls = L64.converted(sloppy)
    bf = gpu.Field(lat, 0, 3, 1, 0).set_lex(b)
    x = gpu.Field(lat, 0, 3, 1, 0)
    info = gpu.stag_cg_mixed(L64, Ls, x, bf, mass, 1e-20, 20000)
    assert info["converged"], info
    xs = x.get_lex()
    assert _true_rel_resid(dims, links, mass, b, xs) <= 1e-20 * 1.01
    xr, _, _ = O.cg(dims, links, mass, b, 1e-24, 20000, noscale=True)
    assert rel_err(xs, xr, parity_mask(dims, 0)) < 1e-7


def test_fb_rejected_where_unsupported(gpu):
    """Forward-only sets run single-rhs operators only: a block of rhs, the
    block and the mixed multishift solvers refuse them with an error; a set
    without long links does not pair with one that has them."""
    dims = (4, 4, 4, 8)
    fl, ll, _ = _fb_case(dims, 3)
    lat = gpu.Lattice(dims)
    L = _fb_links(gpu, lat, fl, ll)
    yb = gpu.Field(lat, 0, 3, 4, 1)
    ob = gpu.Field(lat, 0, 3, 4, 0)
    with pytest.raises(gpu.QllAmdError):
        gpu.dslash(L, ob, yb)
    xb = gpu.Field(lat, 0, 3, 4, 0)
    bb = gpu.Field(lat, 0, 3, 4, 0)
    with pytest.raises(gpu.QllAmdError):
        gpu.stag_cg_mixed_block(L, L.converted(1), xb, bb, 0.1, 1e-10, 10)
    x = gpu.Field(lat, 0, 3, 1, 0)
    bf = gpu.Field(lat, 0, 3, 1, 0)
    with pytest.raises(gpu.QllAmdError):
        gpu.stag_cgms_mixed(L, L.converted(1), [x], bf, 0.1, [0.0], 1e-10, 10)
    Lfat = _fb_links(gpu, lat, fl, None)
    with pytest.raises(gpu.QllAmdError):
        gpu.stag_cg_mixed(L, Lfat.converted(1), x, bf, 0.1, 1e-10, 10)


@pytest.mark.parametrize("sloppy", [1, 3])
def test_mixed_cg_doublestored_fp64_forward_only_sloppy(gpu, sloppy):
    """Doublestored FP64 links with forward-only sloppy links of the same
    operator: the mixed solve reaches the FP64 true residual 1e-10 with the
    oracle's operator."""
    dims = (8, 8, 8, 8)
    mass = 0.05
    fl, ll, _ = _fb_case(dims, 17)
    links = O.stag_doublestore(dims, fl, ll)
    b = O.gauss_vector(dims, 18, 100, parity=0)
    lat = gpu.Lattice(dims)
    L64 = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    Ls = _fb_links(gpu, lat, fl, ll, sloppy)
    bf = gpu.Field(lat, 0, 3, 1, 0).set_lex(b)
    x = gpu.Field(lat, 0, 3, 1, 0)
    info = gpu.stag_cg_mixed(L64, Ls, x, bf, mass, 1e-20, 20000)
    assert info["converged"], info
    assert _true_rel_resid(dims, links, mass, b, x.get_lex()) <= 1e-20 * 1.01
