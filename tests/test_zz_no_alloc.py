"""igr_congestion_view_noalloc (no device allocation: static per-GPU scratch)
gives igr_congestion_view's outputs bit for bit, on calls in a row on one GPU
with a smaller grid in between, and refuses grids above its limit."""
import numpy as np
import pytest

from instantgr_amd import gpu
from instantgr_amd.host import Design

pytestmark = pytest.mark.gpu


def _routed(d):
    r = gpu.Router.from_design(d)
    r.set_nets_from_design(d)
    r.pattern_route(d.array("order"), form=gpu.FORM_BATCHED)
    return r


def test_noalloc_view_matches_the_allocating_one():
    big = Design.preset(0).build(8)                      # 200x200 gcells
    small = Design.custom(layers=6, xsize=48, ysize=40, nets=300, pin_cap=12, sigma=5.0, seed=7).build(4)
    r1 = _routed(big)
    x1, y1 = r1.congestion_view()
    assert np.any(x1 > 0) and np.any(y1 > 0)
    a1, b1 = r1.congestion_view_noalloc()
    assert np.array_equal(a1.view(np.uint32), x1.view(np.uint32))
    assert np.array_equal(b1.view(np.uint32), y1.view(np.uint32))
    # a second call in a row, on a smaller grid (the scratch keeps the first
    # call's larger contents past the smaller grid's end)
    r2 = _routed(small)
    x2, y2 = r2.congestion_view()
    a2, b2 = r2.congestion_view_noalloc()
    assert np.array_equal(a2.view(np.uint32), x2.view(np.uint32))
    assert np.array_equal(b2.view(np.uint32), y2.view(np.uint32))
    # and the first grid again
    a3, b3 = r1.congestion_view_noalloc()
    assert np.array_equal(a3.view(np.uint32), x1.view(np.uint32))
    assert np.array_equal(b3.view(np.uint32), y1.view(np.uint32))
    r1.close()
    r2.close()


def test_noalloc_view_refuses_grids_above_the_limit():
    X = Y = 2900                                          # 8.41M gcells > 2^23
    cap = np.zeros(X * Y, np.float32)
    r = gpu.Router(1, X, Y, 0, 1.0, 4.0, np.ones(1), np.ones(X - 1, np.int32), np.ones(Y - 1, np.int32), cap)
    with pytest.raises(RuntimeError, match="no-allocation limit"):
        r.congestion_view_noalloc()
    xs, _ = r.congestion_view()                           # the allocating entry point still serves it
    assert xs.shape == (X * Y,)
    r.close()
