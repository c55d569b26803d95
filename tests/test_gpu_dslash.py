"""GPU parity: asqtad / staggered and Wilson even-odd Dslash against the FP64 oracle.

Tolerances (relative L2 difference over the output parity):
  FP64 1e-13, FP32 1e-5 (the parity targets), FP16 5e-3 (10-bit mantissa
  storage of links and vectors with per-site max-norm vectors).
"""
import numpy as np
import pytest

import oracle as O
from helpers import parity_mask, rel_err

pytestmark = pytest.mark.gpu

TOL = {0: 1e-13, 1: 1e-5, 2: 5e-3, 3: 2e-4}   # 3: 16-bit fixed point (15-bit mantissas)
DIMS = [(4, 4, 4, 4), (8, 8, 8, 8), (4, 6, 8, 10), (6, 4, 4, 12)]


def _stag_case(dims, seed, nterms=16):
    if nterms == 16:
        links = O.asqtad_links(dims, seed)
    else:
        u = O.apply_stag_phases(dims, O.random_su3_field(dims, seed))
        fl, _ = O.fat7(dims, u, u, want_long=False)
        links = O.stag_doublestore(dims, fl, None)
    y = O.gauss_vector(dims, seed + 1, 100, parity=-1)
    return links, y


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_asqtad_eo_dslash(gpu, dims, prec):
    print("case: asqtad EO dims=%s prec=%d" % (dims, prec), flush=True)
    links, y = _stag_case(dims, 11)
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    if prec != 0:
        L = L.converted(prec)
    for p in (0, 1):
        yin = gpu.Field(lat, prec, 3, 1, 1 - p).set_lex(y)
        out = gpu.Field(lat, prec, 3, 1, p)
        gpu.dslash(L, out, yin)
        ref = O.dslash(dims, links, 0.0, y, sub="even" if p == 0 else "odd")
        err = rel_err(out.get_lex(), ref, parity_mask(dims, p))
        assert err < TOL[prec], (p, err)


@pytest.mark.parametrize("nterms", [8, 16])
def test_dslash_all_with_mass(gpu, nterms):
    """QLL dslash2 on 'all': x = mass*y + D y, both parities (lib/stagDslash.c:313)."""
    dims = (4, 4, 6, 8)
    links, y = _stag_case(dims, 5, nterms)
    mass = 0.37
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.STAG, 0, nterms).set_lex(links)
    ref = O.dslash(dims, links, mass, y, sub="all")
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


def test_normal_operator(gpu):
    """dslashp2: Ap = m^2 p - D_eo D_oe p, pAp = m^2|p|^2 + |D_oe p|^2."""
    dims = (4, 4, 4, 8)
    links, _ = _stag_case(dims, 3)
    p = O.gauss_vector(dims, 9, 101, parity=0)
    mass = 0.2
    Ap_ref, an_ref = O.normal_op(dims, links, mass, p)
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    pf = gpu.Field(lat, 0, 3, 1, 0).set_lex(p)
    Ap = gpu.Field(lat, 0, 3, 1, 0)
    t = gpu.Field(lat, 0, 3, 1, 1)
    an = gpu.stag_normal(L, Ap, pf, t, mass)
    assert abs(an - an_ref) <= 1e-13 * an_ref
    assert rel_err(Ap.get_lex(), Ap_ref, parity_mask(dims, 0)) < 1e-13
    # p^+ A p equals Anorm2 (the identity QLL relies on)
    assert abs(pf.redot(Ap) - an_ref) <= 1e-12 * an_ref


@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_wilson_eo_dslash(gpu, prec):
    dims = (4, 4, 6, 8)
    u = O.random_su3_field(dims, 21)
    links = O.wilson_doublestore(dims, u)
    y = O.gauss_vector(dims, 22, 200, parity=-1, nreal=24)
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.WILSON, 0, 8).set_lex(links)
    if prec != 0:
        L = L.converted(prec)
    mass = 0.15
    ref = O.wilson_dslash(dims, links, mass, y, sub="all")
    for p in (0, 1):
        yo = gpu.Field(lat, prec, 12, 1, 1 - p).set_lex(y)
        yp = gpu.Field(lat, prec, 12, 1, p).set_lex(y)
        out = gpu.Field(lat, prec, 12, 1, p)
        gpu.dslash(L, out, yo, a=4 + mass, w=yp, b=-1.0)
        err = rel_err(out.get_lex(), ref, parity_mask(dims, p))
        assert err < TOL[prec], (p, err)


@pytest.mark.parametrize("nrhs", [4, 16])
def test_wilson_multi_rhs(gpu, nrhs):
    """A block of right-hand sides sharing link loads equals nrhs single solves."""
    dims = (4, 4, 4, 8)
    links = O.wilson_doublestore(dims, O.random_su3_field(dims, 31))
    ys = np.stack([O.gauss_vector(dims, 40 + r, 300, parity=-1, nreal=24) for r in range(nrhs)])
    lat = gpu.Lattice(dims)
    L64 = gpu.Links(lat, gpu.WILSON, 0, 8).set_lex(links)
    L = L64.converted(1)
    yin = gpu.Field(lat, 1, 12, nrhs, 1).set_lex(ys)
    out = gpu.Field(lat, 1, 12, nrhs, 0)
    gpu.dslash(L, out, yin, b=-1.0)   # QLL's Dslash subtracts the hopping sum
    got = out.get_lex()
    m = parity_mask(dims, 0)
    for r in range(nrhs):
        ref = O.wilson_dslash(dims, links, 0.0, ys[r], sub="even")
        err = rel_err(got[r], ref, m)
        assert err < 1e-5, (r, err)


@pytest.mark.parametrize("prec,nrhs", [(0, 3), (0, 8), (1, 5), (1, 16), (2, 16), (3, 7)])
def test_asqtad_block_dslash(gpu, prec, nrhs):
    """The block Dslash (links loaded once for all rhs) gives the oracle's
    Dslash on every rhs, with the a*w + b*D form and both parities."""
    print("case: asqtad block prec=%d nrhs=%d" % (prec, nrhs), flush=True)
    dims = (4, 6, 4, 8)
    links, _ = _stag_case(dims, 21)
    ys = np.stack([O.gauss_vector(dims, 60 + r, 100, parity=-1) for r in range(nrhs)])
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    if prec:
        L = L.converted(prec)
    a, b = 0.3, -0.7
    for p in (0, 1):
        yo = gpu.Field(lat, prec, 3, nrhs, 1 - p).set_lex(ys)
        yp = gpu.Field(lat, prec, 3, nrhs, p).set_lex(ys)
        out = gpu.Field(lat, prec, 3, nrhs, p)
        gpu.dslash(L, out, yo, a=a, w=yp, b=b)
        got = out.get_lex()
        m = parity_mask(dims, p)
        for r in range(nrhs):
            ref = a * ys[r] + b * O.dslash(dims, links, 0.0, ys[r], sub="even" if p == 0 else "odd")
            err = rel_err(got[r], ref, m)
            assert err < TOL[prec], (p, r, err)


@pytest.mark.parametrize("prec,oprec,nrhs", [(3, 3, 8), (3, 1, 5), (2, 2, 4), (2, 1, 8)])
def test_block_dslash_single_input(gpu, prec, oprec, nrhs):
    """The block Dslash with 16-bit links and FP32 input (the block CG's
    D_oe p; the links' term scale applied to the link instead of the vector)
    gives the oracle's Dslash on every rhs, output in the links' format or
    FP32."""
    print("case: block single-in prec=%d out=%d nrhs=%d" % (prec, oprec, nrhs), flush=True)
    dims = (4, 6, 4, 8)
    links, _ = _stag_case(dims, 29)
    ys = np.stack([O.gauss_vector(dims, 80 + r, 100, parity=-1) for r in range(nrhs)])
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links).converted(prec)
    for p in (0, 1):
        yo = gpu.Field(lat, 1, 3, nrhs, 1 - p).set_lex(ys)
        out = gpu.Field(lat, oprec, 3, nrhs, p)
        gpu.dslash(L, out, yo)
        got = out.get_lex()
        m = parity_mask(dims, p)
        for r in range(nrhs):
            ref = O.dslash(dims, links, 0.0, ys[r], sub="even" if p == 0 else "odd")
            err = rel_err(got[r], ref, m)
            assert err < TOL[prec], (p, r, err)


@pytest.mark.parametrize("inner", [(1, 1, 1, 1), (2, 1, 1, 1), (1, 2, 2, 1), (2, 2, 2, 1)])
def test_qll_layout_transfer(gpu, inner):
    """Device import/export of QLL's SIMD layout (lib/layout.c:120-142) for several VLEN."""
    dims = (4, 4, 4, 8)
    lo = O.layout(dims, inner)
    v = O.gauss_vector(dims, 3, 7, parity=-1)
    q = O.to_qll(lo, v)
    import ctypes as C
    import qll_amd
    ql = qll_amd.QllLayout()
    for i in range(4):
        ql.physGeom[i] = lo.physGeom[i]
        ql.innerGeom[i] = lo.innerGeom[i]
        ql.outerGeom[i] = lo.outerGeom[i]
    ql.nSitesOuter, ql.nSitesInner, ql.nEvenOuter = lo.nSitesOuter, lo.nSitesInner, lo.nEvenOuter
    ql.innerCb, ql.innerCbDir = lo.innerCb, lo.innerCbDir
    lat = gpu.Lattice(dims)
    back = np.zeros_like(q)
    for p in (0, 1):
        f = gpu.Field(lat, 0, 3, 1, p)
        qll_amd._check(qll_amd.lib().qamd_field_from_qll(f.h, C.byref(ql), q.ctypes.data, 0))
        assert rel_err(f.get_lex(), v, parity_mask(dims, p)) == 0.0
        qll_amd._check(qll_amd.lib().qamd_field_to_qll(f.h, C.byref(ql), back.ctypes.data, 0))
    assert np.array_equal(back, q)


@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (32, 16, 8, 8)])
def test_block_dslash_every_launch_shape(gpu, dims):
    """Every launch shape the autotuner can pick (rhs per launch, block order:
    plain, XCD ranges, and on (32, 16, 8, 8) the slab sweep) gives the block
    Dslash and the block CG of the default shape bit for bit (partial sums
    are indexed by site block, not by launch order)."""
    print("case: block shapes dims=%s" % (dims,), flush=True)
    links, _ = _stag_case(dims, 23)
    ys = np.stack([O.gauss_vector(dims, 90 + r, 100, parity=1) for r in range(6)])
    bs = np.stack([O.gauss_vector(dims, 95 + r, 100, parity=0) for r in range(6)])
    lat = gpu.Lattice(dims)
    L64 = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    L = L64.converted(1)
    yin = gpu.Field(lat, 1, 3, 6, 1).set_lex(ys)
    b = gpu.Field(lat, 0, 3, 6, 0).set_lex(bs)
    res = []
    for mode in [0] + list(range(2, 11)) + [1]:   # default, each candidate, tuned
        gpu.set_tuning(mode)
        out = gpu.Field(lat, 1, 3, 6, 0)
        gpu.dslash(L, out, yin)
        x = gpu.Field(lat, 0, 3, 6, 0)
        infos = gpu.stag_cg_mixed_block(L64, L, x, b, 0.1, 1e-20, 5000)
        res.append((out.get_lex(), x.get_lex(), [i["iterations"] for i in infos]))
    gpu.set_tuning(1)
    for o, x, its in res[1:]:
        assert np.array_equal(o, res[0][0])
        assert np.array_equal(x, res[0][1]) and its == res[0][2]


@pytest.mark.parametrize("prec,nrhs", [(0, 1), (0, 2), (1, 16), (3, 8)])
def test_wilson_reducing_launches_fixed_tiling(gpu, prec, nrhs):
    """On a lattice that admits the 32x4x4x4 site tiling every reducing Wilson
    launch uses it: all shapes the tuner can pick give the same output and the
    same |out|^2 bit for bit, both matching the oracle's Dslash."""
    print("case: wilson reducing shapes prec=%d nrhs=%d" % (prec, nrhs), flush=True)
    dims = (32, 4, 4, 8)
    links = O.wilson_doublestore(dims, O.random_su3_field(dims, 37))
    ys = np.stack([O.gauss_vector(dims, 90 + r, 300, parity=-1, nreal=24) for r in range(nrhs)])
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.WILSON, 0, 8).set_lex(links)
    if prec:
        L = L.converted(prec)
    yin = gpu.Field(lat, prec, 12, nrhs, 1).set_lex(ys)
    res, norms = [], []
    for mode in [0] + list(range(2, 20)) + [1]:   # default, each candidate, tuned
        gpu.set_tuning(mode)
        out = gpu.Field(lat, prec, 12, nrhs, 0)
        norms.append(gpu.dslash(L, out, yin, b=-1.0, want_norm=True))
        res.append(out.get_lex())
    gpu.set_tuning(1)
    for o in res[1:]:
        assert np.array_equal(o, res[0])
    assert all(n == norms[0] for n in norms)
    got = res[0] if nrhs > 1 else res[0][None]
    m = parity_mask(dims, 0)
    nref = 0.0
    for r in range(nrhs):
        ref = O.wilson_dslash(dims, links, 0.0, ys[r], sub="even")
        assert rel_err(got[r], ref, m) < TOL[prec], r
        nref += float(np.sum(ref[m] ** 2))
    assert abs(norms[0] - nref) / nref < 4 * TOL[prec]


@pytest.mark.parametrize("prec,nrhs,dims", [(0, 1, (8, 8, 8, 8)), (1, 16, (8, 8, 8, 8)), (3, 8, (8, 8, 8, 8)),
                                             (0, 8, (8, 8, 8, 8)), (0, 1, (32, 16, 8, 8))])
def test_wilson_every_launch_shape(gpu, prec, nrhs, dims):
    """Every Wilson launch shape the autotuner can pick (block order, links
    staged in LDS, tiled site orders, slab sweep) gives the same output bit
    for bit, and the oracle's Dslash; with the |out|^2 reduction too.
    32x16x8x8 admits the slab sweep for one rhs (x-y planes of 256 sites)."""
    print("case: wilson shapes prec=%d nrhs=%d dims=%s" % (prec, nrhs, dims), flush=True)
    links = O.wilson_doublestore(dims, O.random_su3_field(dims, 33))
    ys = np.stack([O.gauss_vector(dims, 70 + r, 300, parity=-1, nreal=24) for r in range(nrhs)])
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.WILSON, 0, 8).set_lex(links)
    if prec:
        L = L.converted(prec)
    yin = gpu.Field(lat, prec, 12, nrhs, 1).set_lex(ys)
    res, norms = [], []
    for mode in [0] + list(range(2, 20)) + [1]:   # default, each candidate, tuned
        gpu.set_tuning(mode)
        out = gpu.Field(lat, prec, 12, nrhs, 0)
        gpu.dslash(L, out, yin, b=-1.0)
        res.append(out.get_lex())
        norms.append(gpu.dslash(L, out, yin, b=-1.0, want_norm=True))
    gpu.set_tuning(1)
    for o in res[1:]:
        assert np.array_equal(o, res[0])
    assert all(n == norms[0] for n in norms)
    got = res[0] if nrhs > 1 else res[0][None]
    m = parity_mask(dims, 0)
    for r in range(nrhs):
        ref = O.wilson_dslash(dims, links, 0.0, ys[r], sub="even")
        assert rel_err(got[r], ref, m) < TOL[prec], r


@pytest.mark.parametrize("prec", [0, 1, 3])
def test_mall_share_same_bits(gpu, prec):
    """The MALL share (how many 3-hop link groups load through the caches
    instead of non-temporally) changes cache hints only: every share gives
    the same bits, and they match the oracle."""
    dims = (16, 16, 16, 16)   # one 3-hop group: 6 x 16 B x 32768 sites = 3.1 MB
    print("case: MALL share dims=%s prec=%d" % (dims, prec), flush=True)
    links, y = _stag_case(dims, 23)
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    assert L.reconstructs
    if prec != 0:
        L = L.converted(prec)
    yin = gpu.Field(lat, prec, 3, 1, 1).set_lex(y)
    out = gpu.Field(lat, prec, 3, 1, 0)
    outs = []
    try:
        for mb in (0, 4, 7, 1000):   # 0, 1, 2 and every 3-hop group
            gpu.set_mall_share(mb, mb)
            out.zero()
            gpu.dslash(L, out, yin)
            outs.append(out.get_lex())
    finally:
        gpu.set_mall_share(-1, -1)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    ref = O.dslash(dims, links, 0.0, y, sub="even")
    assert rel_err(outs[0], ref, parity_mask(dims, 0)) < TOL[prec]
