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


@pytest.mark.parametrize("coef", [
    (0.7, 0.0, 0.0, 0.0, 0.0, 0.0),          # one-link term only (no staple pass)
    (0.6, -0.05, 0.0, 0.0, 0.0, 0.0),        # 3-staples only
    (0.6, -0.05, 0.02, 0.0, 0.0, 0.0),       # no 7-staples: the 5-staple sums are not stored
    (0.6, 0.0, 0.0, 0.0, 0.03, 0.0),         # Lepage only
    (0.6, -0.05, 0.02, -0.003, 0.03, -0.04),
])
def test_fat7_coefficient_subsets(gpu, coef):
    """Each staple pass switched on or off by its coefficients (the passes and
    their temporaries are skipped when unused) against the oracle."""
    dims = (4, 6, 4, 8)
    ref_u = O.random_su3_field(dims, 21)
    ref_fl, ref_ll = O.fat7(dims, ref_u, ref_u, coef)
    lat = gpu.Lattice(dims)
    u = gpu.Gauge(lat).set_lex(ref_u)
    fl, ll = gpu.fat7(u, coef)
    assert rel_err(fl.get_lex(), ref_fl) < 1e-13
    if coef[5] != 0:
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


@pytest.mark.parametrize("sloppy", [1, 3])
def test_mixed_cg_keep_and_restart(gpu, sloppy, monkeypatch):
    """Reliable updates that keep the search direction (measured r.s, p.s,
    r.p; the default) and ones that restart it (QAMD_CG_KEEP=0) both reach
    the FP64 true residual and the oracle's solution; keeping takes no more
    iterations and restarts only on a stall."""
    dims = (8, 8, 8, 8)
    mass = 0.05
    links, b = _solver_case(dims, 12)
    lat = gpu.Lattice(dims)
    L64 = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    Ls = L64.converted(sloppy)
    bf = gpu.Field(lat, 0, 3, 1, 0).set_lex(b)
    xr, _, _ = O.cg(dims, links, mass, b, 1e-24, 20000, noscale=True)
    res = {}
    for keep in ("1", "0"):
        monkeypatch.setenv("QAMD_CG_KEEP", keep)
        x = gpu.Field(lat, 0, 3, 1, 0)
        info = gpu.stag_cg_mixed(L64, Ls, x, bf, mass, 1e-20, 20000)
        assert info["converged"] and info["reliable_updates"] >= 2, (keep, info)
        xs = x.get_lex()
        assert _true_rel_resid(dims, links, mass, b, xs) <= 1e-20 * 1.01, keep
        assert rel_err(xs, xr, parity_mask(dims, 0)) < 1e-7, keep
        res[keep] = info
    monkeypatch.delenv("QAMD_CG_KEEP")
    assert res["1"]["restarts"] == 0 and res["0"]["restarts"] == res["0"]["reliable_updates"] - 1
    assert res["1"]["iterations"] <= res["0"]["iterations"]


@pytest.mark.parametrize("nrhs", [1, 3, 8])
def test_fp64_block_cg_matches_oracle(gpu, nrhs):
    """Plain FP64 block CG: every rhs reaches the oracle CG's solution and
    true residual; its iteration count is the oracle's rounded up to the
    16-iteration check (a zero source leaves at once)."""
    dims = (8, 8, 8, 8)
    mass, rsq = 0.2, 1e-16
    links = O.asqtad_links(dims, 3)
    bs = np.stack([O.gauss_vector(dims, 60 + r, 100, parity=0) * (1.0 + 3 * r) for r in range(nrhs)])
    if nrhs == 3:
        bs[1] = 0.0
    lat = gpu.Lattice(dims)
    L = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    bb = gpu.Field(lat, 0, 3, nrhs, 0).set_lex(bs)
    xb = gpu.Field(lat, 0, 3, nrhs, 0)
    infos = gpu.stag_cg_block(L, xb, bb, mass, rsq, 2000)
    xs = xb.get_lex().reshape(nrhs, -1, 6)
    m = parity_mask(dims, 0)
    for r in range(nrhs):
        assert infos[r]["converged"], (r, infos[r])
        if not bs[r].any():
            assert infos[r]["iterations"] == 0 and not xs[r].any()
            continue
        xr, its_r, _ = O.cg(dims, links, mass, bs[r], rsq, 2000, noscale=True)
        assert its_r <= infos[r]["iterations"] <= its_r + 16, (r, infos[r]["iterations"], its_r)
        assert _true_rel_resid(dims, links, mass, bs[r], xs[r]) < 4 * rsq
        assert rel_err(xs[r], xr, m) < 1e-6


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


@pytest.mark.parametrize("sloppy,nrhs,dims", [(None, 1, (4, 4, 4, 8)), (1, 4, (4, 4, 4, 8)), (2, 4, (4, 4, 4, 8)),
                                            (None, 2, (32, 4, 4, 4)), (1, 16, (32, 4, 4, 4)), (3, 8, (32, 4, 4, 4))])
def test_wilson_cgne_block(gpu, sloppy, nrhs, dims):
    """CGNE on Mhat = 1 - k^2 H_eo H_oe for a block of right-hand sides, each
    to FP64 true residual 1e-10 (checked with the oracle's Mhat) and to the
    oracle's own CGNE solution.  The 32 x 4 x 4 x 4 lattice runs the reducing
    Dslash launches (fused residual epilogues) in the tiled site order."""
    print("case: wilson cgne sloppy=%s nrhs=%d dims=%s" % (sloppy, nrhs, dims), flush=True)
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


@pytest.mark.parametrize("dims", [(6, 4, 4, 10), (4, 6, 10, 4)])
def test_fused_solver_steps_on_ragged_lattices(gpu, dims):
    """The fused x/r steps (single-rhs mixed CG, block CG with one mass per
    rhs, multishift refinement) on lattices whose site count per parity is
    not a multiple of the 256-site block: every solution reaches the FP64
    true residual 1e-10 with the oracle's operator."""
    print("case: fused steps dims=%s" % (dims,), flush=True)
    links, b = _solver_case(dims, 40)
    lat = gpu.Lattice(dims)
    L64 = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    LQ = L64.converted(gpu.Q16)
    bf = gpu.Field(lat, 0, 3, 1, 0).set_lex(b)
    mass = 0.1
    x = gpu.Field(lat, 0, 3, 1, 0)
    assert gpu.stag_cg_mixed(L64, LQ, x, bf, mass, 1e-20, 20000)["converged"]
    assert _true_rel_resid(dims, links, mass, b, x.get_lex()) <= 1.01e-20
    bs = np.stack([O.gauss_vector(dims, 50 + r, 100, parity=0) for r in range(3)])
    bb = gpu.Field(lat, 0, 3, 3, 0).set_lex(bs)
    xb = gpu.Field(lat, 0, 3, 3, 0)
    infos = gpu.stag_cg_mixed_block(L64, LQ, xb, bb, mass, 1e-20, 20000)
    xbs = xb.get_lex().reshape(3, -1, 6)
    for r in range(3):
        assert infos[r]["converged"]
        assert _true_rel_resid(dims, links, mass, bs[r], xbs[r]) <= 1.01e-20, r
    shifts = [0.0, 0.01, 0.3]
    xs = [gpu.Field(lat, 0, 3, 1, 0) for _ in shifts]
    infos = gpu.stag_cgms_mixed(L64, LQ, xs, bf, mass, shifts, 1e-20, 20000)
    for i, sh in enumerate(shifts):
        assert infos[i]["converged"]
        assert _true_rel_resid(dims, links, np.sqrt(mass ** 2 + sh), b, xs[i].get_lex()) <= 1.01e-20, i


# ---------------------------------------------------------------------------
# The reference configurations.  Oracle-sized volumes are checked against the
# oracle; the full sizes (no oracle in reach) by FP64 true residuals computed
# with the device's FP64 operator, itself oracle-checked above.

def test_config0_fp64_cg_8x8x8x8(gpu):
    """8^4, seeded SU(3) + fat7/Naik, Gaussian even source, QLL mass 0.2
    (m = 0.1), |r|/|b| = 1e-8: FP64 CG with QLL's recurrences gives the
    oracle's iteration count and solution."""
    dims = (8, 8, 8, 8)
    links = O.asqtad_links(dims, 100)
    b = O.gauss_vector(dims, 101, 100, parity=0)
    xr, its_r, _ = O.cg(dims, links, 0.2, b, 1e-16, 5000, noscale=True)
    lat = gpu.Lattice(dims)
    L = gpu.asqtad_links_device(lat, 100)
    bf = gpu.Field(lat, 0, 3, 1, 0).set_lex(b)
    x = gpu.Field(lat, 0, 3, 1, 0)
    info = gpu.stag_cg(L, x, bf, 0.2, 1e-16, 5000)
    assert info["converged"] and abs(info["iterations"] - its_r) <= 1, (info, its_r)
    xs = x.get_lex()
    assert rel_err(xs, xr, parity_mask(dims, 0)) < 1e-9
    assert _true_rel_resid(dims, links, 0.2, b, xs) < 1.1e-16


def _c2_shifts():
    """12 shifts log-spaced in [1e-4, 1] on top of m^2 = 0.005^2 (QLL mass
    0.01); sigma_0 is folded into the base mass, QLL's cgms taking shifts[0]
    as 0 (lib/cgms.c:6-179)."""
    sig = np.logspace(-4, 0, 12)
    m0 = np.sqrt(0.01 ** 2 + sig[0])
    return sig, m0, list(sig - sig[0])


def test_config2_cgms_12_shifts_oracle_volume(gpu):
    """Config 2 at 8^4: FP32 multishift + FP64 refinement; every shift reaches
    FP64 true |r|/|b| = 1e-10 with the oracle's operator and the oracle's
    FP64 multishift solution."""
    dims = (8, 8, 8, 8)
    links, b = _solver_case(dims, 20)
    sig, m0, shifts = _c2_shifts()
    lat = gpu.Lattice(dims)
    L64 = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    bf = gpu.Field(lat, 0, 3, 1, 0).set_lex(b)
    xs = [gpu.Field(lat, 0, 3, 1, 0) for _ in shifts]
    infos = gpu.stag_cgms_mixed(L64, L64.converted(1), xs, bf, m0, shifts, 1e-20, 50000)
    masses = [np.sqrt(m0 ** 2 + s) for s in shifts]
    xr, _ = O.cgms(dims, links, masses, b, 1e-24, 50000, noscale=True)
    m = parity_mask(dims, 0)
    for i in range(len(shifts)):
        assert infos[i]["converged"], (i, infos[i])
        x = xs[i].get_lex()
        assert _true_rel_resid(dims, links, masses[i], b, x) <= 1.01e-20, i
        assert rel_err(x, xr[i], m) < 1e-8, i


@pytest.mark.parametrize("sloppy", [1, 2, 3])
def test_config3_wilson_16rhs_oracle_volume(gpu, sloppy):
    """Config 3 at 4^3x8: 16 Gaussian spinor sources, kappa = 0.124, sloppy
    FP32/FP16/Q16 Dslash sharing link loads across the 16 rhs, FP64 defect
    correction; every rhs at FP64 true |r|/|b| = 1e-10 (oracle's Mhat)."""
    dims = (4, 4, 4, 8)
    links = O.wilson_doublestore(dims, O.random_su3_field(dims, 43))
    bs = np.stack([O.gauss_vector(dims, 200 + r, 400, parity=0, nreal=24) for r in range(16)])
    lat = gpu.Lattice(dims)
    L64 = gpu.Links(lat, gpu.WILSON, 0, 8).set_lex(links)
    b = gpu.Field(lat, 0, 12, 16, 0).set_lex(bs)
    x = gpu.Field(lat, 0, 12, 16, 0)
    infos = gpu.wilson_cgne(L64, x, b, 0.124, 1e-20, 5000, ls=L64.converted(sloppy))
    xs = x.get_lex().reshape(16, -1, 24)
    m = parity_mask(dims, 0)
    for r in range(16):
        assert infos[r]["converged"], (r, infos[r])
        res = np.sum((bs[r] - O.wilson_meo(dims, links, 0.124, xs[r]))[m] ** 2) / np.sum(bs[r][m] ** 2)
        assert res <= 1.01e-20, (r, res)


def _device_true_rsq(gpu, L64, x, b, mass):
    lat = L64.lat
    Ap = gpu.Field(lat, 0, 3, x.nrhs, 0)
    t = gpu.Field(lat, 0, 3, x.nrhs, 1)
    gpu.stag_normal(L64, Ap, x, t, mass)
    Ap.axpby(1.0, b, -1.0)
    r2, b2 = Ap.norm2(), b.norm2()
    if x.nrhs == 1:
        return r2 / b2
    return max(r2[i] / b2[i] for i in range(x.nrhs))


def test_config1_full_32x32x32x32_mixed_vs_fp64(gpu):
    """Config 1 at full size: FP64 CG and the 16-bit mixed CG both reach FP64
    true |r|/|b| <= 1e-10 and agree on the solution."""
    dims = (32, 32, 32, 32)
    print("case: config1 full 32^4", flush=True)
    lat = gpu.Lattice(dims)
    L64 = gpu.asqtad_links_device(lat, 1)
    LQ = L64.converted(gpu.Q16)
    b = gpu.Field(lat, 0, 3, 1, 0).gaussian(11, 100)
    x64 = gpu.Field(lat, 0, 3, 1, 0)
    xm = gpu.Field(lat, 0, 3, 1, 0)
    assert gpu.stag_cg(L64, x64, b, 0.02, 1e-20, 50000)["converged"]
    assert gpu.stag_cg_mixed(L64, LQ, xm, b, 0.02, 1e-20, 50000)["converged"]
    assert _device_true_rsq(gpu, L64, xm, b, 0.02) <= 1.01e-20
    assert _device_true_rsq(gpu, L64, x64, b, 0.02) <= 1.01e-20
    d = gpu.Field(lat, 0, 3, 1, 0).copy_from(xm)
    d.axpby(-1.0, x64, 1.0)
    assert (d.norm2() / x64.norm2()) ** 0.5 < 1e-6


def test_config2_cgms_q16_refinement_groups(gpu):
    """12 shifts at 8^4 with Q16 sloppy links (base m^2 = 1e-4, lighter than
    config 2): the shifts are refined in blocks of up to 8 rhs, lightest
    first, one mass per rhs; a block holding a shift whose mass is too light
    for the 16-bit links (operator error >= 0.9 m^2) runs on FP32 links.
    Every shift reaches FP64 true |r|/|b| = 1e-10 with the oracle's operator."""
    dims = (8, 8, 8, 8)
    links, b = _solver_case(dims, 21)
    m0 = 0.01
    shifts = list(np.logspace(-4, 0, 12) - 1e-4)
    lat = gpu.Lattice(dims)
    L64 = gpu.Links(lat, gpu.STAG, 0, 16).set_lex(links)
    LQ = L64.converted(gpu.Q16)
    eps = gpu.stag_sloppy_error(L64, LQ, m0)
    bf = gpu.Field(lat, 0, 3, 1, 0).set_lex(b)
    xs = [gpu.Field(lat, 0, 3, 1, 0) for _ in shifts]
    infos = gpu.stag_cgms_mixed(L64, LQ, xs, bf, m0, shifts, 1e-20, 50000)
    blocks = [list(range(0, 8)), list(range(8, 12))]   # shifts are ascending
    want = {}
    for blk in blocks:
        coarse = any(eps >= 0.9 * (m0 ** 2 + shifts[i]) for i in blk)
        want.update({i: int(coarse) for i in blk})
    assert [i["escalations"] for i in infos] == [want[i] for i in range(12)], (eps, infos)
    assert want[0] == 1 and want[11] == 0, eps                # both link precisions are used
    for i, s in enumerate(shifts):
        assert infos[i]["converged"], (i, infos[i])
        x = xs[i].get_lex()
        assert _true_rel_resid(dims, links, np.sqrt(m0 ** 2 + s), b, x) <= 1.01e-20, i


def test_config2_full_32x32x32x64_cgms(gpu):
    """Config 2 at full size: 12 shifts, FP32 inner multishift + FP64
    refinement; every shift at FP64 true |r|/|b| <= 1e-10."""
    dims = (32, 32, 32, 64)
    print("case: config2 full 32^3x64", flush=True)
    sig, m0, shifts = _c2_shifts()
    lat = gpu.Lattice(dims)
    L64 = gpu.asqtad_links_device(lat, 2)
    b = gpu.Field(lat, 0, 3, 1, 0).gaussian(12, 100)
    xs = [gpu.Field(lat, 0, 3, 1, 0) for _ in shifts]
    infos = gpu.stag_cgms_mixed(L64, L64.converted(gpu.F32), xs, b, m0, shifts, 1e-20, 50000)
    for i, s in enumerate(shifts):
        assert infos[i]["converged"], (i, infos[i])
        assert _device_true_rsq(gpu, L64, xs[i], b, float(np.sqrt(m0 ** 2 + s))) <= 1.01e-20, i


def test_config3_full_32x32x32x64_wilson_16rhs(gpu):
    """Config 3 at full size: 16 rhs, kappa 0.124, FP32 inner; every rhs at
    FP64 true |r|/|b| <= 1e-10.  The residual is recomputed outside the
    solver, b - Mhat x with Mhat x = x - k^2 H_eo H_oe x from two FP64 Wilson
    Dslash launches, and for two rhs with the oracle's Mhat on the host."""
    dims = (32, 32, 32, 64)
    kappa = 0.124
    print("case: config3 full 32^3x64 x 16 rhs", flush=True)
    lat = gpu.Lattice(dims)
    g = gpu.Gauge(lat).random_su3(3)
    L64 = gpu.Links(lat, gpu.WILSON, 0, 8).from_gauge(g)
    b = gpu.Field(lat, 0, 12, 16, 0).gaussian(13, 100)
    x = gpu.Field(lat, 0, 12, 16, 0)
    infos = gpu.wilson_cgne(L64, x, b, kappa, 1e-20, 5000, ls=L64.converted(gpu.F32))
    for r in range(16):
        assert infos[r]["converged"], (r, infos[r])
    # t = -H_oe x, res = -x - k^2 (-H_eo) t + b = b - Mhat x   (Dslash b=-1 is -H)
    t = gpu.Field(lat, 0, 12, 16, 1)
    gpu.dslash(L64, t, x, b=-1.0)
    res = gpu.Field(lat, 0, 12, 16, 0)
    gpu.dslash(L64, res, t, a=-1.0, w=x, b=-kappa * kappa)
    res.axpby(1.0, b, 1.0)
    r2, b2 = res.norm2(), b.norm2()
    for r in range(16):
        assert r2[r] / b2[r] <= 1.01e-20, (r, r2[r] / b2[r], infos[r])
    del t, res
    links = O.wilson_doublestore(dims, g.get_lex())
    xs = x.get_lex()
    bs = b.get_lex()
    m = parity_mask(dims, 0)
    for r in (0, 15):
        Mx = O.wilson_meo(dims, links, kappa, xs[r])
        rr = np.sum((bs[r] - Mx)[m] ** 2) / np.sum(bs[r][m] ** 2)
        assert rr <= 1.01e-20, (r, rr)
