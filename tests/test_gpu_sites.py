"""GPU: transfers in a caller's site order (the order a QDP/QOP code exposes,
given as a coordinate table; QDP's row-major matrices transposed), checked
against the lexicographic transfers, plus the table's error checks."""
import numpy as np
import pytest

import oracle as O
from helpers import parity_mask

pytestmark = pytest.mark.gpu


def _table(dims, seed):
    """Every site once, in a random order (as QDP_get_coords_L would list them)."""
    V = O.volume(dims)
    lex = np.random.default_rng(seed).permutation(V)
    c = np.empty((V, 4), dtype=np.int32)
    rem = lex.copy()
    for d in range(4):
        c[:, d] = rem % dims[d]
        rem //= dims[d]
    return c, lex


@pytest.mark.parametrize("ncomp,hdtype", [(3, np.float64), (3, np.float32), (12, np.float64)])
def test_field_site_table_matches_lex(gpu, ncomp, hdtype):
    dims = (4, 6, 4, 8)
    coords, lex = _table(dims, 3)
    y = O.gauss_vector(dims, 9, 100, parity=-1, nreal=2 * ncomp)
    lat = gpu.Lattice(dims)
    ysite = y[lex].astype(hdtype)
    for p in (0, 1):
        a = gpu.Field(lat, 0, ncomp, 1, p).set_sites(coords, ysite)
        b = gpu.Field(lat, 0, ncomp, 1, p).set_lex(y.astype(hdtype).astype(np.float64))
        assert np.array_equal(a.get_lex(), b.get_lex())
        base = np.full((O.volume(dims), 2 * ncomp), 7.0)
        back = a.get_sites(coords, base=base)
        mine = parity_mask(dims, p)[lex]
        assert np.array_equal(back[mine], ysite[mine].astype(np.float64))
        assert np.all(back[~mine] == 7.0)   # the other parity is left as it was


def test_gauge_site_table_row_major(gpu):
    dims = (4, 4, 4, 6)
    coords, lex = _table(dims, 5)
    u = O.random_su3_field(dims, 11)                      # (4, V, 18) column-major
    ucm = u.reshape(4, -1, 3, 3, 2)                       # [mu][site][c][r][re/im]
    urm = np.ascontiguousarray(ucm.transpose(0, 1, 3, 2, 4)).reshape(4, -1, 18)   # row-major
    lat = gpu.Lattice(dims)
    g = gpu.Gauge(lat).set_sites(coords, urm[:, lex], row_major=True)
    assert np.array_equal(g.get_lex(), u)
    assert np.array_equal(g.get_sites(coords, row_major=True), urm[:, lex])
    g2 = gpu.Gauge(lat).set_sites(coords, u[:, lex])      # column-major, as QLL
    assert np.array_equal(g2.get_lex(), u)


def test_dslash_on_site_ordered_inputs(gpu):
    """Links made from a site-ordered gauge field and a site-ordered source give
    the same Dslash, bit for bit, as the lexicographic path."""
    dims = (4, 4, 6, 4)
    coords, lex = _table(dims, 8)
    u = O.random_su3_field(dims, 12)
    y = O.gauss_vector(dims, 13, 100, parity=1)
    lat = gpu.Lattice(dims)
    outs = []
    for site_order in (False, True):
        g = gpu.Gauge(lat)
        if site_order:
            g.set_sites(coords, u[:, lex])
        else:
            g.set_lex(u)
        L = gpu.Links(lat, gpu.STAG, 0, 8).from_gauge(g)
        yin = gpu.Field(lat, 0, 3, 1, 1)
        if site_order:
            yin.set_sites(coords, y[lex])
        else:
            yin.set_lex(y)
        out = gpu.Field(lat, 0, 3, 1, 0)
        gpu.dslash(L, out, yin)
        outs.append(out.get_sites(coords) if site_order else out.get_lex()[lex])
    m = parity_mask(dims, 0)[lex]
    assert np.array_equal(outs[0][m], outs[1][m])


def test_site_table_errors(gpu):
    dims = (4, 4, 4, 4)
    coords, lex = _table(dims, 1)
    lat = gpu.Lattice(dims)
    f = gpu.Field(lat, 0, 3, 1, 0)
    y = np.zeros((O.volume(dims), 6))
    bad = coords.copy()
    bad[5, 2] = 4
    with pytest.raises(gpu.QllAmdError, match="outside"):
        f.set_sites(bad, y)
    dup = coords.copy()
    dup[7] = dup[3]
    with pytest.raises(gpu.QllAmdError, match="same coordinates"):
        f.set_sites(dup, y)
    even = parity_mask(dims, 0)[lex]
    short = coords[even][1:]                              # one even site missing
    with pytest.raises(gpu.QllAmdError, match="even and"):
        f.set_sites(short, y[: short.shape[0]])
    # the odd sites alone are not enough for an even field, but enough for an odd one
    odd = coords[~even]
    gpu.Field(lat, 0, 3, 1, 1).set_sites(odd, y[: odd.shape[0]])
    with pytest.raises(gpu.QllAmdError, match="nrhs = 1"):
        gpu.Field(lat, 0, 3, 2, 0).set_sites(coords, y)
