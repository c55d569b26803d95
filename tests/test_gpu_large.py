"""GPU: the largest lattice of the suite, 64^4 (8.4M sites; 38 GB of FP64
asqtad links), checked through identities that hold at any size: the
staggered operator is anti-hermitian between the parities, linear, and its
FP32 form agrees with FP64; a 16-rhs Wilson block equals 16 single-rhs
launches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = (64, 64, 64, 64)


@pytest.fixture(scope="module")
def big(gpu):
    lat = gpu.Lattice(DIMS)
    L64 = gpu.asqtad_links_device(lat, 7)
    return lat, L64


def test_asqtad_identities_at_64_4(gpu, big):
    print("case: 64^4 asqtad identities", flush=True)
    lat, L64 = big
    xe = gpu.Field(lat, 0, 3, 1, 0).gaussian(1, 100)
    yo = gpu.Field(lat, 0, 3, 1, 1).gaussian(2, 100)
    Dy = gpu.Field(lat, 0, 3, 1, 0)
    Dx = gpu.Field(lat, 0, 3, 1, 1)
    gpu.dslash(L64, Dy, yo)
    gpu.dslash(L64, Dx, xe)
    # <x_e, D_eo y_o> = -conj(<D_oe x_e, y_o>): the real parts cancel
    a, b = xe.redot(Dy), Dx.redot(yo)
    scale = (xe.norm2() * Dy.norm2()) ** 0.5
    assert abs(a + b) <= 1e-12 * scale, (a, b)
    # linearity: D(2 y + z) = 2 D y + D z
    zo = gpu.Field(lat, 0, 3, 1, 1).gaussian(3, 100)
    Dz = gpu.Field(lat, 0, 3, 1, 0)
    gpu.dslash(L64, Dz, zo)
    w = gpu.Field(lat, 0, 3, 1, 1).copy_from(zo).axpby(2.0, yo, 1.0)
    Dw = gpu.Field(lat, 0, 3, 1, 0)
    gpu.dslash(L64, Dw, w)
    Dw.axpby(-2.0, Dy, 1.0).axpby(-1.0, Dz, 1.0)
    assert Dw.norm2() <= 1e-26 * Dy.norm2()
    # FP32 links and vectors: FP32-level agreement
    L32 = L64.converted(gpu.F32)
    y32 = gpu.Field(lat, 1, 3, 1, 1).copy_from(yo)
    D32 = gpu.Field(lat, 1, 3, 1, 0)
    gpu.dslash(L32, D32, y32)
    d = gpu.Field(lat, 0, 3, 1, 0).copy_from(D32).axpby(-1.0, Dy, 1.0)
    assert d.norm2() <= (1e-6) ** 2 * Dy.norm2()


def test_wilson_16rhs_block_at_64_4(gpu, big):
    print("case: 64^4 wilson 16-rhs block", flush=True)
    lat, _ = big
    g = gpu.Gauge(lat).random_su3(9)
    W = gpu.Links(lat, gpu.WILSON, 0, 8).from_gauge(g)
    del g
    W32 = W.converted(gpu.F32)
    del W
    blk = gpu.Field(lat, 1, 12, 16, 1).gaussian(4, 100)
    out = gpu.Field(lat, 1, 12, 16, 0)
    gpu.dslash(W32, out, blk, b=-1.0)
    n_blk = out.norm2()
    for r in (0, 7, 15):   # rhs r of the block draws Gaussian stream 100 + r
        one = gpu.Field(lat, 1, 12, 1, 1).gaussian(4, 100 + r)
        o1 = gpu.Field(lat, 1, 12, 1, 0)
        gpu.dslash(W32, o1, one, b=-1.0)
        # the single-rhs launch rebuilds the links' third rows from the FP32
        # first two, the 16-rhs block reads the stored FP32 third rows: both
        # are FP32 roundings of the same SU(3) matrices
        assert n_blk[r] == pytest.approx(o1.norm2(), rel=1e-6), r
