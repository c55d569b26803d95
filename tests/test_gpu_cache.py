"""GPU: reuse of freed device blocks (8 MB and more) and their release."""
import pytest

pytestmark = pytest.mark.gpu


def test_freed_blocks_reused_zeroed_and_released(gpu):
    """A field created in a block another field left behind starts at zero,
    like a fresh one, and the cached blocks can be released at any time."""
    print("case: device block cache", flush=True)
    lat = gpu.Lattice((32, 32, 32, 32))
    f = gpu.Field(lat, 0, 3, 8, 0).gaussian(1, 100)   # 8 x 3 x 16 B x 524288 sites = 201 MB
    assert min(f.norm2()) > 0
    del f
    g = gpu.Field(lat, 0, 3, 8, 0)                    # the same size: the freed block again
    assert g.norm2() == [0.0] * 8
    g.gaussian(2, 100)
    ref = g.norm2()
    gpu.release_cache()
    h = gpu.Field(lat, 0, 3, 8, 0)
    assert h.norm2() == [0.0] * 8
    assert g.norm2() == ref                           # live fields are not touched
    del g, h
    gpu.release_cache()
