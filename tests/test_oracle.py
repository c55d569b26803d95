"""Pins of the FP64 oracle (CPU only).

The reference stores no golden vectors and its tests need QOPQDP (absent), so
each oracle function is pinned by what the mathematics fixes: closed forms,
invariants and dense brute force on small lattices, each chosen so a plausible
mistake (wrong hop, sign, adjoint, index order, missing term) fails.
"""
import itertools
import os

import numpy as np
import pytest

import oracle as O
from helpers import cplx, mats, parity_mask, rel_err, unmats


def lex(dims, x):
    return x[0] + dims[0] * (x[1] + dims[1] * (x[2] + dims[2] * x[3]))


# ------------------------------------------------------------------ SU(3)
def test_su3_generator_unitary_det1():
    dims = (4, 4, 4, 6)
    M = mats(O.random_su3_field(dims, 3))
    eye = np.eye(3)
    assert np.abs(M @ M.conj().swapaxes(-1, -2) - eye).max() < 1e-14
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-14


def test_su3_generator_unitary_at_scale():
    """Over the 262,144 links of a 16^4 field the worst matrix stays unitary
    to a few eps.  One Gram-Schmidt pass leaves eps * cond(G) errors, which
    reach 1e-13..1e-12 at this count and break the 1e-12 check that lets the
    FP64 Dslash rebuild the long links' third rows."""
    M = mats(O.random_su3_field((16, 16, 16, 16), 3))
    eye = np.eye(3)
    assert np.abs(M @ M.conj().swapaxes(-1, -2) - eye).max() < 1e-14


def test_su3_gram_schmidt_order():
    """makeU order (lib/vla.c:1635-1654): column 0 of U is the normalised
    column 0 of the Gaussian matrix, times the det phase factor."""
    g = np.random.default_rng(1).standard_normal(18)
    m = np.empty(18)
    O.lib().orc_su3_from_gauss(m, g)
    G = mats(g[None])[0]
    U = mats(m[None])[0]
    c0 = G[:, 0] / np.linalg.norm(G[:, 0])
    ph = U[0, 0] / c0[0]
    assert abs(abs(ph) - 1) < 1e-14
    assert np.allclose(U[:, 0], ph * c0, atol=1e-14)
    # column 1 is orthogonal to column 0 and lies in span(G0, G1)
    assert abs(np.vdot(U[:, 0], U[:, 1])) < 1e-14
    Q, _ = np.linalg.qr(G[:, :2])
    assert np.linalg.norm(U[:, 1] - Q @ (Q.conj().T @ U[:, 1])) < 1e-13


def test_rng_layout_independent_and_gaussian():
    v = O.gauss_vector((8, 8, 8, 8), 5, 100, parity=-1)
    z = cplx(v).ravel()
    assert abs(np.mean(np.abs(z) ** 2) - 1) < 0.03
    assert abs(np.mean(z.real)) < 0.02
    # value at a site depends only on (seed, stream, global site)
    w = O.gauss_vector((8, 8, 8, 8), 5, 100, parity=0)
    m = parity_mask((8, 8, 8, 8), 0)
    assert np.array_equal(w[m], v[m]) and not w[~m].any()


# ----------------------------------------------------------------- gammas
def test_gamma_algebra_degrand_rossi():
    g = [O.gamma(mu) for mu in range(4)]
    for a in range(4):
        assert np.allclose(g[a], g[a].conj().T)
        for b in range(4):
            assert np.allclose(g[a] @ g[b] + g[b] @ g[a], 2 * (a == b) * np.eye(4))
    g5 = O.gamma(4)
    assert np.allclose(g5, g[0] @ g[1] @ g[2] @ g[3])
    assert np.allclose(g5, np.diag([1, 1, -1, -1]))


# ------------------------------------------------------- staggered Dslash
def _unit_links_with_phases(dims, c_one, c_naik):
    V = O.volume(dims)
    u = np.zeros((4, V, 18))
    u[:, :, 0] = u[:, :, 8] = u[:, :, 16] = 1.0
    u = O.apply_stag_phases(dims, u)
    return O.stag_doublestore(dims, c_one * u, c_naik * u)


def test_free_field_dispersion():
    """Unit gauge with staggered phases: -D^2 on a plane wave e^{ipx} equals
    sum_mu 4 (c1 sin p_mu + c3 sin 3p_mu)^2.  Pins hop distances (1 and 3),
    the -U^+ backward convention, the Naik sign and the phases."""
    dims = (8, 8, 8, 16)
    c1, c3 = 1.2, -1.0 / 24
    links = _unit_links_with_phases(dims, c1, c3)
    X = O.coords(dims)
    for n in [(1, 0, 0, 0), (0, 2, 0, 3), (3, 1, 2, 5)]:
        p = 2 * np.pi * np.array(n) / np.array(dims)
        ph = np.exp(1j * (X @ p))
        col = np.array([0.3 + 0.1j, -0.2j, 0.5])
        z = ph[:, None] * col[None, :]
        v = np.empty((len(ph), 6))
        v[:, 0::2], v[:, 1::2] = z.real, z.imag
        D2 = O.dslash(dims, links, 0.0, O.dslash(dims, links, 0.0, v))
        lam = np.sum(4 * (c1 * np.sin(p) + c3 * np.sin(3 * p)) ** 2)
        assert rel_err(-D2, lam * v) < 1e-12


def test_dslash_antihermitian_and_eo_blocks():
    dims = (4, 4, 4, 6)
    links = O.asqtad_links(dims, 2)
    a = O.gauss_vector(dims, 1, 1, parity=-1)
    b = O.gauss_vector(dims, 1, 2, parity=-1)
    ip = lambda x, y: np.vdot(cplx(x).ravel(), cplx(y).ravel())
    assert abs(ip(a, O.dslash(dims, links, 0, b)) + ip(O.dslash(dims, links, 0, a), b)) < 1e-11
    # the EO blocks: even-subset output only touches even sites, from odd input
    ye = O.gauss_vector(dims, 1, 3, parity=0)
    assert not O.dslash(dims, links, 0, ye, sub="even").any()
    mass = 0.7
    full = O.dslash(dims, links, mass, b, sub="all")
    parts = O.dslash(dims, links, 0, b, sub="even") + O.dslash(dims, links, 0, b, sub="odd")
    assert rel_err(full, parts + mass * b) < 1e-14


@pytest.mark.parametrize("naik", [False, True])
def test_dslash_fb_matches_doublestored(naik):
    """dslashFb2T on forward links equals dslash2 on the doublestored set built
    from them (t[2mu+1] = -F(x-mu)^+, lib/gauge.c:34-63), as testdslash.c:338-351
    checks both against one qopDslash; antihermitian; a backward term with
    the wrong sign or without the adjoint is caught."""
    dims = (4, 6, 4, 8)
    u = O.apply_stag_phases(dims, O.random_su3_field(dims, 4))
    fl, ll = O.fat7(dims, u, u, want_long=naik)
    y = O.gauss_vector(dims, 3, 7, parity=-1)
    links = O.stag_doublestore(dims, fl, ll)
    for sub, mass in (("all", 0.3), ("even", 0.0), ("odd", 0.0)):
        ref = O.dslash(dims, links, mass, y, sub=sub)
        got = O.dslash_fb(dims, fl, ll, mass, y, sub=sub)
        assert rel_err(got, ref) < 1e-14, sub
    ip = lambda x, z: np.vdot(cplx(x).ravel(), cplx(z).ravel())
    b = O.gauss_vector(dims, 3, 8, parity=-1)
    assert abs(ip(y, O.dslash_fb(dims, fl, ll, 0, b)) + ip(O.dslash_fb(dims, fl, ll, 0, y), b)) < 1e-11
    # planted mistakes in the comparison set: +F(x-mu)^+ and -F(x-mu)
    nt = links.shape[0]
    bwd = [t for t in range(nt) if ((t // 2 if nt == 16 else t) & 1)]
    wrong_sign = links.copy()
    wrong_sign[bwd] *= -1
    no_adj = links.copy()
    for t in bwd:   # transpose instead of conjugate transpose
        no_adj[t] = unmats(mats(links[t]).swapaxes(-1, -2))
    ref = O.dslash_fb(dims, fl, ll, 0.3, y)
    for bad in (wrong_sign, no_adj):
        assert rel_err(O.dslash(dims, bad, 0.3, y), ref) > 1e-2


def _gauge_transform(dims, links16, g):
    """L_t(x) -> g(x) L_t(x) g(x + d_t)^+ for the 16-term set."""
    V = O.volume(dims)
    M = mats(links16)
    out = np.empty_like(M)
    for t in range(16):
        k, naik = t // 2, t & 1
        mu, hop = k // 2, (3 if naik else 1)
        d = -hop if (k & 1) else hop
        nb = np.array([O.lib().orc_neighbor((__import__('ctypes').c_int * 4)(*dims), s, mu, d) for s in range(V)])
        out[t] = g @ M[t] @ g[nb].conj().swapaxes(-1, -2)
    return unmats(out)


def test_dslash_gauge_covariance():
    dims = (4, 4, 4, 4)
    links = O.asqtad_links(dims, 6)
    g = mats(O.random_su3_field(dims, 9)[0])
    lt = _gauge_transform(dims, links, g)
    y = O.gauss_vector(dims, 2, 5, parity=-1)
    gy = np.einsum("sab,sb->sa", g, cplx(y))
    gyr = np.empty_like(y)
    gyr[:, 0::2], gyr[:, 1::2] = gy.real, gy.imag
    lhs = cplx(O.dslash(dims, lt, 0.3, gyr))
    rhs = np.einsum("sab,sb->sa", g, cplx(O.dslash(dims, links, 0.3, y)))
    assert np.abs(lhs - rhs).max() < 1e-12


# ----------------------------------------------------------------- fat7
def test_fat7_unit_gauge_closed_form():
    """U = eta (unit gauge with phases): every staple picks up a -1, so
    fat = (c1 - 6c3 + 24c5 - 48c7 + 6cL) eta and long = naik * eta."""
    dims = (4, 4, 4, 4)
    V = O.volume(dims)
    u = np.zeros((4, V, 18))
    u[:, :, 0] = u[:, :, 8] = u[:, :, 16] = 1.0
    u = O.apply_stag_phases(dims, u)
    for coef in [O.FAT7_ASQTAD, (0.3 - 0.1 * 15 / 16, 0.2 / 24, 0.2 / 384, 0.2 / 3072, 0.1 / 96, 0.2)]:
        fl, ll = O.fat7(dims, u, u, coef)
        c1, c3, c5, c7, cL, naik = coef
        ceff = c1 - 6 * c3 + 24 * c5 - 48 * c7 + 6 * cL
        assert np.abs(fl - ceff * u).max() < 1e-14
        assert np.abs(ll - naik * u).max() < 1e-14


def _shift(a, mu, h, dims):
    """a(x + h mu) for a site-major array in lexicographic order"""
    s = a.reshape(tuple(reversed(dims)) + a.shape[1:])
    return np.roll(s, -h, axis=3 - mu).reshape(a.shape)


def _fat7_paths(dims, U, coef):
    """Independent path enumeration: explicit products of links along every
    staple path (no genStaple recursion)."""
    c1, c3, c5, c7, cL, naik = coef
    V = O.volume(dims)
    eye = np.broadcast_to(np.eye(3), (V, 3, 3))

    def link(mu, sgn, at):   # matrix for one step from site 'at' along sgn*mu
        if sgn > 0:
            return at(U[mu])
        return at(_shift(U[mu], mu, -1, dims)).conj().swapaxes(-1, -2)

    def path(steps):   # product along a path starting at x
        P = eye.copy()
        off = np.zeros(4, dtype=int)
        for mu, sgn in steps:
            o = off.copy()
            at = lambda F, o=o: _shiftv(F, o)
            P = P @ link(mu, sgn, at)
            off[mu] += sgn
        return P

    def _shiftv(F, o):
        for mu in range(4):
            if o[mu]:
                F = _shift(F, mu, o[mu], dims)
        return F

    fat = np.zeros((4, V, 3, 3), dtype=complex)
    for mu in range(4):
        fat[mu] = (c1 - 6 * cL) * U[mu]
        others = [d for d in range(4) if d != mu]
        for nu in others:
            for s in (1, -1):
                fat[mu] += c3 * path([(nu, s), (mu, 1), (nu, -s)])
                for s2 in (1, -1):
                    fat[mu] += cL * path([(nu, s2), (nu, s), (mu, 1), (nu, -s), (nu, -s2)])
            for rho in others:
                if rho == nu:
                    continue
                for s in (1, -1):
                    for s2 in (1, -1):
                        fat[mu] += c5 * path([(rho, s2), (nu, s), (mu, 1), (nu, -s), (rho, -s2)])
                        for sig in others:
                            if sig in (nu, rho):
                                continue
                            for s3 in (1, -1):
                                fat[mu] += c7 * path([(sig, s3), (rho, s2), (nu, s), (mu, 1),
                                                      (nu, -s), (rho, -s2), (sig, -s3)])
    lng = np.stack([naik * U[mu] @ _shift(U[mu], mu, 1, dims) @ _shift(U[mu], mu, 2, dims) for mu in range(4)])
    return fat, lng


def test_fat7_brute_force_paths():
    dims = (4, 4, 4, 4)
    u = O.random_su3_field(dims, 17)
    coef = (0.3 - 0.1 * 15 / 16, 0.2 / 24, 0.2 / 384, 0.2 / 3072, 0.1 / 96, 0.2)
    fl, ll = O.fat7(dims, u, u, coef)
    fat, lng = _fat7_paths(dims, mats(u), coef)
    assert np.abs(mats(fl) - fat).max() < 1e-13
    assert np.abs(mats(ll) - lng).max() < 1e-13


def test_fat7_gauge_covariance():
    dims = (4, 4, 4, 6)
    u = O.random_su3_field(dims, 4)
    g = mats(O.random_su3_field(dims, 5)[0])
    U = mats(u)
    Ug = np.stack([g @ U[mu] @ _shift(g, mu, 1, dims).conj().swapaxes(-1, -2) for mu in range(4)])
    fl, ll = O.fat7(dims, u, u)
    flg, llg = O.fat7(dims, unmats(Ug), unmats(Ug))
    for mu in range(4):
        rhs = g @ mats(fl)[mu] @ _shift(g, mu, 1, dims).conj().swapaxes(-1, -2)
        assert np.abs(mats(flg)[mu] - rhs).max() < 1e-13
        rhs = g @ mats(ll)[mu] @ _shift(g, mu, 3, dims).conj().swapaxes(-1, -2)
        assert np.abs(mats(llg)[mu] - rhs).max() < 1e-13


def test_plaquette_unit_and_covariant():
    dims = (4, 4, 4, 4)
    V = O.volume(dims)
    u = np.zeros((4, V, 18))
    u[:, :, 0] = u[:, :, 8] = u[:, :, 16] = 1.0
    assert np.allclose(O.plaquette(dims, u), 1.0)
    r = O.random_su3_field(dims, 1)
    g = mats(O.random_su3_field(dims, 2)[0])
    Ug = unmats(np.stack([g @ mats(r)[mu] @ _shift(g, mu, 1, dims).conj().swapaxes(-1, -2) for mu in range(4)]))
    assert np.allclose(O.plaquette(dims, r), O.plaquette(dims, Ug), atol=1e-14)


# ---------------------------------------------------------------- solvers
def _dense_normal(dims, links, mass):
    """Dense matrix of A = m^2 - D_eo D_oe on the even sites (complex)."""
    ev = np.where(parity_mask(dims, 0))[0]
    n = 3 * len(ev)
    A = np.empty((n, n), dtype=complex)
    for j in range(n):
        e = np.zeros((O.volume(dims), 6))
        e[ev[j // 3], 2 * (j % 3)] = 1.0
        Ae, _ = O.normal_op(dims, links, mass, e)
        A[:, j] = cplx(Ae[ev]).ravel()
    return A, ev


def test_cg_dense_solve():
    dims = (4, 4, 4, 4)
    links = O.asqtad_links(dims, 8)
    mass = 0.3
    b = O.gauss_vector(dims, 2, 100, parity=0)
    A, ev = _dense_normal(dims, links, mass)
    assert np.allclose(A, A.conj().T, atol=1e-13)
    assert np.linalg.eigvalsh(A).min() > mass ** 2 * 0.999
    xd = np.linalg.solve(A, cplx(b[ev]).ravel())
    x, its, fr = O.cg(dims, links, mass, b, 1e-28, 500, noscale=True)
    assert fr < 1e-28
    assert np.abs(cplx(x[ev]).ravel() - xd).max() < 1e-12 * np.abs(xd).max()
    # solve2 scaling: x <- mass x, and (m + D) x_full = b for x_o = -D_oe x_e / m
    xs, _, _ = O.cg(dims, links, mass, b, 1e-28, 500, noscale=False)
    assert rel_err(xs, mass * x) < 1e-15
    xo = -O.dslash(dims, links, 0.0, xs, sub="odd") / mass
    full = xs * parity_mask(dims, 0)[:, None] + xo
    assert rel_err(O.dslash(dims, links, mass, full, sub="all"), b) < 1e-12


def test_cg_iteration_semantics():
    """do-while: at least one iteration, stop at its >= maxits (lib/cg.c:115)."""
    dims = (4, 4, 4, 4)
    links = O.asqtad_links(dims, 8)
    b = O.gauss_vector(dims, 2, 100, parity=0)
    _, its, _ = O.cg(dims, links, 0.3, b, 1.0, 500, noscale=True)
    assert its == 1
    _, its, _ = O.cg(dims, links, 0.3, b, 1e-40, 5, noscale=True)
    assert its == 5


def test_cgms_dense_solves():
    dims = (4, 4, 4, 4)
    links = O.asqtad_links(dims, 10)
    masses = [0.3, 0.42, 0.7, 1.3]
    b = O.gauss_vector(dims, 4, 100, parity=0)
    xs, its = O.cgms(dims, links, masses, b, 1e-28, 1000, noscale=True)
    A0, ev = _dense_normal(dims, links, masses[0])
    rhs = cplx(b[ev]).ravel()
    for i, m in enumerate(masses):
        A = A0 + (m * m - masses[0] ** 2) * np.eye(A0.shape[0])
        xd = np.linalg.solve(A, rhs)
        assert np.abs(cplx(xs[i][ev]).ravel() - xd).max() < 1e-11 * np.abs(xd).max(), i
    # zero source returns zero at once (lib/cgms.c:25-30)
    z, its0 = O.cgms(dims, links, masses, np.zeros_like(b), 1e-16, 10, noscale=True)
    assert its0 == 0 and not any(v.any() for v in z)


# ----------------------------------------------------------------- Wilson
def test_wilson_free_field():
    """U = 1: x = (4+m) y - sum_mu [(1-g) y(x+mu) + (1+g) y(x-mu)] on a plane
    wave is M(p) = (4+m) - sum_mu (2 cos p - 2 i g_mu sin p)."""
    dims = (4, 4, 6, 8)
    V = O.volume(dims)
    u = np.zeros((4, V, 18))
    u[:, :, 0] = u[:, :, 8] = u[:, :, 16] = 1.0
    links = O.wilson_doublestore(dims, u)
    X = O.coords(dims)
    g = [O.gamma(mu) for mu in range(4)]
    rng = np.random.default_rng(0)
    for n in [(1, 0, 0, 0), (0, 1, 2, 3)]:
        p = 2 * np.pi * np.array(n) / np.array(dims)
        spin = rng.standard_normal(4) + 1j * rng.standard_normal(4)
        col = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        ph = np.exp(1j * (X @ p))
        z = ph[:, None, None] * spin[None, :, None] * col[None, None, :]
        y = np.empty((V, 24))
        y[:, 0::2], y[:, 1::2] = z.reshape(V, 12).real, z.reshape(V, 12).imag
        mass = 0.1
        x = O.wilson_dslash(dims, links, mass, y)
        Mp = (4 + mass) * np.eye(4) - sum(2 * np.cos(p[m]) * np.eye(4) - 2j * np.sin(p[m]) * g[m] for m in range(4))
        ref = ph[:, None, None] * (Mp @ spin)[None, :, None] * col[None, None, :]
        assert np.abs(cplx(x).reshape(V, 4, 3) - ref).max() < 1e-12


def test_wilson_gamma5_hermiticity():
    dims = (4, 4, 4, 4)
    links = O.wilson_doublestore(dims, O.random_su3_field(dims, 3))
    a = O.gauss_vector(dims, 1, 7, parity=-1, nreal=24)
    b = O.gauss_vector(dims, 1, 8, parity=-1, nreal=24)
    g5 = np.array([1, 1, -1, -1], dtype=float).repeat(6)
    ip = lambda x, y: np.vdot(cplx(x).ravel(), cplx(y).ravel())
    m = 0.2
    lhs = ip(a, O.wilson_dslash(dims, links, m, b))
    rhs = ip(g5 * O.wilson_dslash(dims, links, m, g5 * a), b)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def test_wilson_meo_and_cgne():
    dims = (4, 4, 4, 4)
    links = O.wilson_doublestore(dims, O.random_su3_field(dims, 3))
    kappa = 0.124
    b = O.gauss_vector(dims, 5, 9, parity=0, nreal=24)
    ev = parity_mask(dims, 0)
    # Mhat = 1 - kappa^2 H_eo H_oe through the QLL Dslash subsets
    t = -O.wilson_dslash(dims, links, 0, b, sub="odd")
    Hb = -O.wilson_dslash(dims, links, 0, t, sub="even")
    ref = b - kappa ** 2 * Hb
    assert rel_err(O.wilson_meo(dims, links, kappa, b), ref, ev) < 1e-14
    # dagger = g5 Mhat g5 is the adjoint
    a = O.gauss_vector(dims, 6, 9, parity=0, nreal=24)
    ip = lambda x, y: np.vdot(cplx(x[ev]).ravel(), cplx(y[ev]).ravel())
    assert abs(ip(a, O.wilson_meo(dims, links, kappa, b)) -
               ip(O.wilson_meo(dims, links, kappa, a, dagger=True), b)) < 1e-12
    x, its, fr = O.wilson_cgne(dims, links, kappa, b, 1e-20, 2000)
    assert fr <= 1e-20
    assert rel_err(O.wilson_meo(dims, links, kappa, x), b, ev) < 1.1e-10


# ----------------------------------------------------------------- layout
def _inner_for(dims, vlen):
    """partitionGeom(og, ig, lg, VLEN, nd, 1) (lib/stagDslash.c:751-786)"""
    nd = len(dims)
    nx, lx = [1] * nd, list(dims)
    ww = vlen
    while ww > 1:
        k = nd - 1
        while lx[k] & 1:
            k -= 1
        for i in range(k - 1, -1, -1):
            if (lx[i] & 1) == 0 and (nx[i] < nx[k] or (nx[i] == nx[k] and lx[i] > lx[k])):
                k = i
        nx[k] *= 2
        lx[k] //= 2
        ww //= 2
    return tuple(nx)


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (8, 4, 6, 4), (4, 8, 8, 12)])
@pytest.mark.parametrize("vlen", [1, 2, 4, 8])
def test_layout_roundtrip(dims, vlen):
    inner = _inner_for(dims, vlen)
    try:
        lo = O.layout(dims, inner)
    except ValueError:
        pytest.skip("QLL rejects this geometry")
    V = O.volume(dims)
    seen = set()
    for x in itertools.product(*[range(d) for d in dims]):
        i = O.layout_index(lo, list(x))
        assert 0 <= i < V and i not in seen
        seen.add(i)
        assert O.layout_coord(lo, i) == list(x)    # lib/layout.c:195-204
        # even sites first; all lanes of an outer site share its parity
        assert (i // lo.nSitesInner < lo.nEvenOuter) == (sum(x) % 2 == 0)
    v = O.gauss_vector(dims, 1, 1, parity=-1)
    assert np.array_equal(O.from_qll(lo, O.to_qll(lo, v), 6), v)


def test_layout_vlen1_is_checkerboarded_lexicographic():
    dims = (4, 4, 4, 6)
    lo = O.layout(dims, (1, 1, 1, 1))
    for x in itertools.product(*[range(d) for d in dims]):
        lx = lex(dims, x)
        want = lx // 2 + (O.volume(dims) // 2 if sum(x) % 2 else 0)
        assert O.layout_index(lo, list(x)) == want


# ------------------------------------------------------------ qop* layer
def test_qop_libraries_export_every_declared_entry_point():
    """oracle/qop.h declares QLL's qop* test entry points (include/common.h:66-86);
    both per-precision builds export all of them."""
    import re
    import subprocess
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    O.build()
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(here, "qop.h")).read(), flags=re.S)
    declared = set(re.findall(r"\b(qop\w+)\s*\(", txt))
    assert len(declared) == 17   # 9 staggered + 8 Wilson
    for prec in "DF":
        out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(here, "libqop%s.so" % prec)],
                             capture_output=True, text=True, check=True).stdout
        exported = set(re.findall(r"\bT (qop\w+)", out))
        assert declared <= exported, declared - exported
