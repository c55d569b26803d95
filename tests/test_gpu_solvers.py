"""GPU parity: gauge generation, fat7/Naik, CG, multishift CG and the
mixed-precision solvers against the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
from helpers import parity_mask, rel_err

pytestmark = pytest.mark.gpu


def test_random_su3_matches_oracle(gpu):
    dims = (4, 4, 6, 8)
    lat = gpu.Lattice(dims)
    u = gpu.Gauge(lat).random_su3(1234).get_lex()
    ref = O.random_su3_field(dims, 1234)
    assert rel_err(u, ref) < 1e-13


def test_phases_fat7_longlinks_match_oracle(gpu):
    dims = (4, 4, 6, 8)
    ref_u = O.apply_stag_phases(dims, O.random_su3_field(dims, 77))
    ref_fl, ref_ll = O.fat7(dims, ref_u, ref_u)
    lat = gpu.Lattice(dims)
    u = gpu.Gauge(lat).random_su3(77).stag_phases()
    assert rel_err(u.get_lex(), ref_u) < 1e-13
    # smear the oracle's own links so the comparison sees fat7 alone
    u.set_lex(ref_u)
    fl, ll = gpu.fat7(u)
    assert rel_err(fl.get_lex(), ref_fl) < 1e-13
    assert rel_err(ll.get_lex(), ref_ll) < 1e-13


def test_fat7_with_lepage(gpu):
    dims = (4, 4, 4, 6)
    coef = (0.3 - 0.1 * 15 / 16, 0.2 / 24, 0.2 / 384, 0.2 / 3072, 0.1 / 96, 0.2)   # testdslash.c:306-312
    ref_u = O.random_su3_field(dims, 5)
    ref_fl, ref_ll = O.fat7(dims, ref_u, ref_u, coef)
    lat = gpu.Lattice(dims)
    u = gpu.Gauge(lat).set_lex(ref_u)
    fl, ll = gpu.fat7(u, coef)
    assert rel_err(fl.get_lex(), ref_fl) < 1e-13
    assert rel_err(ll.get_lex(), ref_ll) < 1e-13


def test_plaquette(gpu):
    dims = (4, 4, 4, 6)
    ref_u = O.random_su3_field(dims, 8)
    lat = gpu.Lattice(dims)
    p = gpu.Gauge(lat).set_lex(ref_u).plaquette()
    assert np.allclose(p, O.plaquette(dims, ref_u), rtol=0, atol=1e-14)


def test_device_links_match_oracle_chain(gpu):
    """SU(3) -> phases -> fat7 -> doublestore entirely on the device gives the
    oracle's Dslash."""
    dims = (4, 4, 4, 8)
    links = O.asqtad_links(dims, 99)
    lat = gpu.Lattice(dims)
    L = gpu.asqtad_links_device(lat, 99)
    y = O.gauss_vector(dims, 4, 100, parity=1)
    yin = gpu.Field(lat, 0, 3, 1, 1).set_lex(y)
    out = gpu.Field(lat, 0, 3, 1, 0)
    gpu.dslash(L, out, yin)
    ref = O.dslash(dims, links, 0.0, y, sub="even")
    assert rel_err(out.get_lex(), ref, parity_mask(dims, 0)) < 1e-12


def _solver_case(dims, seed):
    links = O.asqtad_links(dims, seed)
    b = O.gauss_vector(dims, seed + 1, 100, parity=0)
    return links, b


def _true_rel_resid(dims, links, mass, b, x):
    Ax, _ = O.normal_op(dims, links, mass, x)
    m = parity_mask(dims, 0)
    return np.sum((b - Ax)[m] ** 2) / np.sum(b[m] ** 2)


@pytest.mark.parametrize("prec", [0, 1])
def test_cg_matches_oracle(gpu, prec):
    """cgSolveT recurrences (lib/cg.c:43-122): same iteration count, same solution."""
    dims = (8, 8, 8, 8)
    mass = 0.2
    links, b = _solver_case(dims, 3)
    rsq = 1e-16 if prec == 0 else 1e-10
    xr, its_r, _ = O.cg(dims, links, mass, b, rsq, 2000, noscale=True)
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    if prec:
        L = L.converted(prec)
    bf = gpu.Field(lat, prec, 3, 1, 0).set_lex(b)
    x = gpu.Field(lat, prec, 3, 1, 0)
    info = gpu.stag_cg(L, x, bf, mass, rsq, 2000)
    assert info["converged"]
    if prec == 0:
        assert abs(info["iterations"] - its_r) <= 1, (info, its_r)
        xs = x.get_lex()
        assert _true_rel_resid(dims, links, mass, b, xs) < 4 * rsq
        assert rel_err(xs, xr, parity_mask(dims, 0)) < 1e-6
    else:
        assert abs(info["iterations"] - its_r) <= max(3, its_r // 20), (info, its_r)
        assert rel_err(x.get_lex(), xr, parity_mask(dims, 0)) < 1e-3


def test_cg_maxits_and_zero_source(gpu):
    dims = (4, 4, 4, 4)
    links, b = _solver_case(dims, 4)
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    bf = gpu.Field(lat, 0, 3, 1, 0).set_lex(b)
    x = gpu.Field(lat, 0, 3, 1, 0)
    info = gpu.stag_cg(L, x, bf, 0.05, 1e-30, 7)
    assert info["iterations"] == 7 and not info["converged"]
    _, its_r, _ = O.cg(dims, links, 0.05, b, 1e-30, 7, noscale=True)
    assert its_r == 7
    z = gpu.Field(lat, 0, 3, 1, 0)
    info = gpu.stag_cg(L, x, z, 0.05, 1e-16, 100)
    assert x.norm2() == 0.0


def test_cgms_matches_oracle(gpu):
    """cgmsSolveT (lib/cgms.c:6-179) with solveMulti2's shifts (lib/stagDslash.c:634-639)."""
    dims = (8, 8, 8, 8)
    links, b = _solver_case(dims, 6)
    masses = [0.2 * 1.4 ** i for i in range(5)]
    rsq = 1e-16
    xr, its_r = O.cgms(dims, links, masses, b, rsq, 2000, noscale=True)
    shifts = [m * m - masses[0] ** 2 for m in masses]
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    bf = gpu.Field(lat, 0, 3, 1, 0).set_lex(b)
    xs = [gpu.Field(lat, 0, 3, 1, 0) for _ in masses]
    info = gpu.stag_cgms(L, xs, bf, masses[0], shifts, rsq, 2000)
    assert abs(info["iterations"] - its_r) <= 1
    m = parity_mask(dims, 0)
    for i, mi in enumerate(masses):
        assert rel_err(xs[i].get_lex(), xr[i], m) < 1e-6
        assert _true_rel_resid(dims, links, mi, b, xs[i].get_lex()) < 1e-14


@pytest.mark.parametrize("sloppy", [1, 2, 3])
def test_mixed_cg_true_residual(gpu, sloppy):
    """FP32/FP16 inner Dslash with FP64 reliable updates reaches the FP64
    true residual 1e-10 and the oracle's FP64 solution."""
    dims = (8, 8, 8, 8)
    mass = 0.05
    links, b = _solver_case(dims, 12)
    lat = gpu.Lattice(dims)
    L64 = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    Ls = L64.converted(sloppy)
    bf = gpu.Field(lat, 0, 3, 1, 0).set_lex(b)
    x = gpu.Field(lat, 0, 3, 1, 0)
    info = gpu.stag_cg_mixed(L64, Ls, x, bf, mass, 1e-20, 20000)
    assert info["converged"], info
    xs = x.get_lex()
    assert _true_rel_resid(dims, links, mass, b, xs) <= 1e-20 * 1.01
    xr, _, _ = O.cg(dims, links, mass, b, 1e-24, 20000, noscale=True)
    assert rel_err(xs, xr, parity_mask(dims, 0)) < 1e-7


@pytest.mark.parametrize("sloppy,nrhs", [(1, 1), (1, 5), (3, 8)])
def test_mixed_cg_block(gpu, sloppy, nrhs):
    """Batched mixed-precision CG: every rhs of the block reaches the FP64
    true residual 1e-10 (checked with the oracle's operator) and the oracle's
    FP64 solution; sources of very different size leave the block at their
    own iteration."""
    print("case: block cg sloppy=%d nrhs=%d" % (sloppy, nrhs), flush=True)
    dims = (8, 8, 8, 8)
    mass = 0.05
    links = O.asqtad_links(dims, 14)
    bs = np.stack([O.gauss_vector(dims, 70 + r, 100, parity=0) * (10.0 ** (r % 3)) for r in range(nrhs)])
    lat = gpu.Lattice(dims)
    L64 = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    Ls = L64.converted(sloppy)
    b = gpu.Field(lat, 0, 3, nrhs, 0).set_lex(bs)
    x = gpu.Field(lat, 0, 3, nrhs, 0)
    infos = gpu.stag_cg_mixed_block(L64, Ls, x, b, mass, 1e-20, 20000)
    xs = x.get_lex().reshape(nrhs, -1, 6)
    m = parity_mask(dims, 0)
    for r in range(nrhs):
        assert infos[r]["converged"], (r, infos[r])
        assert _true_rel_resid(dims, links, mass, bs[r], xs[r]) <= 1e-20 * 1.01, r
    xr, _, _ = O.cg(dims, links, mass, bs[0], 1e-24, 20000, noscale=True)
    assert rel_err(xs[0], xr, m) < 1e-7


def test_mixed_cg_block_zero_rhs(gpu):
    """A zero source in the block gives x = 0 and does not hold up the rest."""
    dims = (4, 4, 4, 8)
    links = O.asqtad_links(dims, 15)
    bs = np.stack([O.gauss_vector(dims, 80, 100, parity=0), np.zeros((O.volume(dims), 6))])
    lat = gpu.Lattice(dims)
    L64 = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    b = gpu.Field(lat, 0, 3, 2, 0).set_lex(bs)
    x = gpu.Field(lat, 0, 3, 2, 0)
    infos = gpu.stag_cg_mixed_block(L64, L64.converted(1), x, b, 0.1, 1e-20, 5000)
    xs = x.get_lex()
    assert infos[0]["converged"] and infos[1]["converged"]
    assert np.all(xs[1] == 0)
    assert _true_rel_resid(dims, links, 0.1, bs[0], xs[0]) <= 1.01e-20


@pytest.mark.parametrize("sloppy,nrhs", [(None, 1), (1, 4), (2, 4)])
def test_wilson_cgne_block(gpu, sloppy, nrhs):
    """CGNE on Mhat = 1 - k^2 H_eo H_oe for a block of right-hand sides, each
    to FP64 true residual 1e-10 (checked with the oracle's Mhat) and to the
    oracle's own CGNE solution."""
    dims = (4, 4, 4, 8)
    kappa = 0.124
    links = O.wilson_doublestore(dims, O.random_su3_field(dims, 41))
    bs = np.stack([O.gauss_vector(dims, 50 + r, 400, parity=0, nreal=24) for r in range(nrhs)])
    lat = gpu.Lattice(dims)
    L64 = gpu.Links(lat, gpu.WILSON, 0, 8).set_lex(links)
    Ls = L64.converted(sloppy) if sloppy else None
    b = gpu.Field(lat, 0, 12, nrhs, 0).set_lex(bs)
    x = gpu.Field(lat, 0, 12, nrhs, 0)
    infos = gpu.wilson_cgne(L64, x, b, kappa, 1e-20, 5000, ls=Ls)
    xs = x.get_lex().reshape(nrhs, -1, 24)
    m = parity_mask(dims, 0)
    for r in range(nrhs):
        assert infos[r]["converged"], infos[r]
        Mx = O.wilson_meo(dims, links, kappa, xs[r])
        res = np.sum((bs[r] - Mx)[m] ** 2) / np.sum(bs[r][m] ** 2)
        assert res <= 1.01e-20, (r, res)
    xr, _, _ = O.wilson_cgne(dims, links, kappa, bs[0], 1e-24, 5000)
    assert rel_err(xs[0], xr, m) < 1e-8


def test_mixed_cgms_all_shifts(gpu):
    dims = (8, 8, 8, 8)
    links, b = _solver_case(dims, 13)
    mass = 0.02
    shifts = list(np.logspace(-4, 0, 6))
    lat = gpu.Lattice(dims)
    L64 = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    L32 = L64.converted(1)
    bf = gpu.Field(lat, 0, 3, 1, 0).set_lex(b)
    xs = [gpu.Field(lat, 0, 3, 1, 0) for _ in shifts]
    infos = gpu.stag_cgms_mixed(L64, L32, xs, bf, mass, shifts, 1e-20, 20000)
    for i, s in enumerate(shifts):
        meff = np.sqrt(mass ** 2 + s)
        assert infos[i]["converged"]
        assert _true_rel_resid(dims, links, meff, b, xs[i].get_lex()) <= 1.01e-20
