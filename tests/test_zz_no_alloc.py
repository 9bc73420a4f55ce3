"""lsg_expand_noalloc (the Expand entry point that allocates nothing) against
lsg_expand through the C ABI: bit-identical run lists on two calls in a row,
the second on a smaller level set, and the oracle's Expand on the first;
refusals that leave the level set unchanged."""
import ctypes

import numpy as np
import pytest

import cases
import pyoracle as po
from test_gpu_parity import assert_same, say, to_gpu

pytestmark = pytest.mark.gpu


def _lib():
    from viennals_amd._lib import lib
    return lib()


def _err():
    return _lib().lsg_last_error().decode()


def _expand_both(o, width):
    """The same level set expanded by lsg_expand and by lsg_expand_noalloc."""
    a, b = to_gpu(o), to_gpu(o)
    assert _lib().lsg_expand(a._h, width) == 0, _err()
    rc = _lib().lsg_expand_noalloc(b._h, width)
    assert rc == 0, _err()
    ia, va = a.get_runs()
    ib, vb = b.get_runs()
    assert np.array_equal(ia, ib)
    assert np.array_equal(va.view(np.int64), vb.view(np.int64))
    assert b.getLevelSetWidth() == width
    return b


def test_noalloc_matches_expand_on_two_calls():
    say("no-alloc expand: trench 96 x 96, width 2 -> 5")
    big = cases.trench(L=48, width=20, depth=18)
    g = _expand_both(big, 5)
    o = big.copy()
    o.expand(5)
    assert_same(o, g, "no-alloc expand of the trench")
    say("no-alloc expand: sphere r = 6, width 2 -> 4 (smaller input)")
    small = cases.sphere(r=6.0, dx=0.5)
    g = _expand_both(small, 4)
    o = small.copy()
    o.expand(4)
    assert_same(o, g, "no-alloc expand of the sphere")


def test_noalloc_2d_and_repeated_calls():
    say("no-alloc expand: circle, width 1 -> 3 -> 6 on the same level set")
    c = cases.sphere2d(r=5.0)
    c.reduce(1)
    g = to_gpu(c)
    for w in (3, 6):
        assert _lib().lsg_expand_noalloc(g._h, w) == 0, _err()
        c.expand(w)
        assert_same(c, g, f"no-alloc expand of the circle to width {w}")
    # a width at or below the current one changes nothing
    i0, v0 = g.get_runs()
    assert _lib().lsg_expand_noalloc(g._h, 4) == 0
    i1, v1 = g.get_runs()
    assert np.array_equal(i0, i1) and np.array_equal(v0.view(np.int64), v1.view(np.int64))


def test_noalloc_refusals_leave_the_level_set_unchanged():
    import viennals_amd as vls
    say("no-alloc expand: band too close to the frame edge")
    t = cases.trench(L=24, width=10, depth=8)
    g = to_gpu(t)
    i0, v0 = g.get_runs()
    assert _lib().lsg_expand_noalloc(g._h, 40) == -44
    i1, v1 = g.get_runs()
    assert np.array_equal(i0, i1) and np.array_equal(v0.view(np.int64), v1.view(np.int64))
    assert g.getLevelSetWidth() == t.width
    say("no-alloc expand: frame above 2^22 words")
    n = 8192  # 257 words x 16385 rows
    d = vls.Domain.from_index_bounds(2, [-n, -n], [n, n], [cases.REF, cases.REF], 1.0)
    vls.MakeGeometry(d, vls.Sphere([0.0, 0.0], 5.0)).apply()
    i0, v0 = d.get_runs()
    w0 = d.getLevelSetWidth()
    assert _lib().lsg_expand_noalloc(d._h, w0 + 2) == -49
    assert "exceeds the limit" in _err()
    i1, v1 = d.get_runs()
    assert np.array_equal(i0, i1) and np.array_equal(v0.view(np.int64), v1.view(np.int64))
    # the allocating entry point still takes it
    assert _lib().lsg_expand(d._h, w0 + 2) == 0, _err()
