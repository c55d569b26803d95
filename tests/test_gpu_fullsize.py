"""GPU parity at the benchmark sizes against the FP64 oracle run on the host.

The oracle builds the 32^4 asqtad links itself (seeded SU(3) -> staggered
phases -> fat7/Naik -> doublestore, oracle.c), so these tests pin the device
Dslash at the size the throughput is quoted on, independently of the device
fat7:

  * asqtad EO Dslash, both parities, at FP64 / FP32 / FP16 / Q16 against
    orc_dslash (lib/stagDslash.c:259-420 semantics), relative L2 error
    1e-13 / 1e-5 / 5e-3 / 2e-4;
  * device SU(3) generation and fat7/Naik at 32^4 against the oracle's;
  * the Wilson FP32 16-rhs block Dslash at 32^3x64 against
    orc_wilson_dslash for every rhs (1e-5).

Host memory: the 32^4 FP64 link set is 2.4 GB as a lexicographic array.
"""
import numpy as np
import pytest

import oracle as O
from helpers import parity_mask, rel_err

pytestmark = pytest.mark.gpu

C1 = (32, 32, 32, 32)
C3 = (32, 32, 32, 64)
TOL = {0: 1e-13, 1: 1e-5, 2: 5e-3, 3: 2e-4}


@pytest.fixture(scope="module")
def c1_oracle():
    print("case: building 32^4 oracle links (host fat7)", flush=True)
    u = O.apply_stag_phases(C1, O.random_su3_field(C1, 1))
    fl, ll = O.fat7(C1, u, u)
    links = O.stag_doublestore(C1, fl, ll)
    y = O.gauss_vector(C1, 2, 100, parity=-1)
    ref = {p: O.dslash(C1, links, 0.0, y, sub="even" if p == 0 else "odd") for p in (0, 1)}
    return u, fl, ll, links, y, ref


@pytest.fixture(scope="module")
def c1_device(gpu, c1_oracle):
    _, _, _, links, _, _ = c1_oracle
    lat = gpu.Lattice(C1)
    L64 = gpu.Links(lat, gpu.STAG, gpu.F64, 16).set_lex(links)
    return lat, L64


@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_asqtad_eo_dslash_32x32x32x32(gpu, c1_oracle, c1_device, prec):
    print("case: 32^4 asqtad EO dslash prec=%d" % prec, flush=True)
    _, _, _, _, y, ref = c1_oracle
    lat, L64 = c1_device
    assert L64.reconstructs          # the benchmarked (two-row long link) kernel
    L = L64 if prec == 0 else L64.converted(prec)
    for p in (0, 1):
        yin = gpu.Field(lat, prec, 3, 1, 1 - p).set_lex(y)
        out = gpu.Field(lat, prec, 3, 1, p)
        gpu.dslash(L, out, yin)
        err = rel_err(out.get_lex(), ref[p], parity_mask(C1, p))
        print("  parity %d rel err %.3e" % (p, err), flush=True)
        assert err < TOL[prec], (p, err)


def test_asqtad_block_dslash_32x32x32x32_fp32(gpu, c1_oracle, c1_device):
    """The 8-rhs block kernel at 32^4 in FP32: every rhs is the scaled input
    plus the oracle's Dslash (a*w + b*D y)."""
    print("case: 32^4 asqtad block dslash fp32 8 rhs", flush=True)
    _, _, _, links, y, ref = c1_oracle
    lat, L64 = c1_device
    L = L64.converted(gpu.F32)
    ys = np.stack([y * (1.0 + 0.25 * r) for r in range(8)])
    yo = gpu.Field(lat, 1, 3, 8, 1).set_lex(ys)
    ye = gpu.Field(lat, 1, 3, 8, 0).set_lex(ys)
    out = gpu.Field(lat, 1, 3, 8, 0)
    gpu.dslash(L, out, yo, a=0.5, w=ye, b=-1.0)
    got = out.get_lex()
    m = parity_mask(C1, 0)
    for r in range(8):
        want = 0.5 * ys[r] - (1.0 + 0.25 * r) * ref[0]
        assert rel_err(got[r], want, m) < TOL[1], r


def test_su3_and_fat7_32x32x32x32(gpu, c1_oracle):
    """Device SU(3) links with phases, and device fat7/Naik smearing of the
    oracle's links, against the oracle at 32^4."""
    print("case: 32^4 su3 + fat7", flush=True)
    u_ref, fl_ref, ll_ref, _, _, _ = c1_oracle
    lat = gpu.Lattice(C1)
    u = gpu.Gauge(lat).random_su3(1).stag_phases()
    assert rel_err(u.get_lex(), u_ref) < 1e-13
    u.set_lex(u_ref)
    fl, ll = gpu.fat7(u)
    assert rel_err(fl.get_lex(), fl_ref) < 1e-13
    assert rel_err(ll.get_lex(), ll_ref) < 1e-13


def test_wilson_fp32_16rhs_32x32x32x64(gpu):
    """Config 3's Dslash: 16 rhs sharing FP32 link loads at 32^3x64, every rhs
    against the oracle's Wilson Dslash."""
    print("case: 32^3x64 wilson fp32 16 rhs", flush=True)
    links = O.wilson_doublestore(C3, O.random_su3_field(C3, 3))
    lat = gpu.Lattice(C3)
    L = gpu.Links(lat, gpu.WILSON, gpu.F64, 8).set_lex(links).converted(gpu.F32)
    m = parity_mask(C3, 0)
    ys = np.empty((16, lat.V, 24))
    for r in range(16):
        ys[r] = O.gauss_vector(C3, 50 + r, 100, parity=1, nreal=24)
    yin = gpu.Field(lat, 1, 12, 16, 1).set_lex(ys)
    out = gpu.Field(lat, 1, 12, 16, 0)
    gpu.dslash(L, out, yin, b=-1.0)
    got = out.get_lex()
    del out, yin
    for r in range(16):
        ref = O.wilson_dslash(C3, links, 0.0, ys[r], sub="even")
        err = rel_err(got[r], ref, m)
        assert err < TOL[1], (r, err)
