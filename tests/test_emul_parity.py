"""Host build of the device per-item logic (tests/emul, compiled from
viennals_amd/csrc/lsg_items.hpp) against the CPU oracle: the bitmap frame,
rank lookups, candidate dilation, compaction and index rebuild must give the
oracle's canonical run list bit for bit.  No GPU needed."""
import numpy as np
import pytest

import cases
import pyemul as pe
import pyoracle as po


@pytest.fixture(autouse=True, params=["tiled", "lists", "pairs"])
def pipeline(request):
    """Every case through the band-operation pipelines of the device code:
    the single-pass tiles, the candidate lists, and the candidate lists with
    Expand cycles two at a time (OP_EXPAND2) and each stage's update fused
    with its rebuild (OP_UPDATE_REBUILD)."""
    pe.set_tiled(request.param == "tiled")
    pe.set_expand_pairs(request.param == "pairs")
    yield request.param
    pe.set_tiled(True)
    pe.set_expand_pairs(False)


def to_emul(o):
    e = pe.Domain(o.D, o.gmin, o.gmax, o.bc, o.delta)
    e.set(*o.get(), o.width)
    return e


def assert_same(o, e, tag):
    io, vo = o.get()
    ie, ve = e.get()
    if not cases.bit_equal(io, vo, ie, ve):
        pytest.fail(f"{tag}: differs at {cases.first_diff(io, vo, ie, ve)}")


def test_band_ops_sphere():
    o = cases.sphere(r=6.0, dx=0.25)
    e = to_emul(o)
    assert_same(o, e, "upload")
    for w in (3, 5, 7, 9):
        o.expand(w)
        e.expand(w)
        assert_same(o, e, f"expand {w}")
    o.reduce(1)
    e.reduce(1)
    assert_same(o, e, "reduce")
    o.rebuild()
    e.rebuild()
    assert_same(o, e, "rebuild")
    for rsz in (True, False):
        o.prune(rsz)
        e.prune(rsz)
        assert_same(o, e, f"prune {rsz}")


@pytest.mark.parametrize("op", [0, 1, 2])
def test_boolean_ops(op):
    L = 24
    gmin, gmax, bc = [-L, -L, -L], [L, L, L], [0, 0, 1]
    a = po.Domain(3, gmin, gmax, bc, 1.0)
    ip, vp = cases.plane_points(3, gmin, gmax, bc, 1.0, 0.0)
    a.from_points(ip, vp, 2)
    b = po.Domain(3, gmin, gmax, bc, 1.0)
    ib, vb = cases.box_points(3, [-6, -6, -10], [6, 6, 1], 1.0, gmin, gmax)
    b.from_points(ib, vb, 2)
    ea, eb = to_emul(a), to_emul(b)
    a.boolean(b, op)
    ea.boolean(eb, op)
    assert_same(a, ea, f"boolean {op}")


GEOMS = {
    "sphere": lambda: cases.sphere(r=5.0, dx=0.25, origin=[0.3, 0.1, -0.2]),
    "trench": lambda: cases.trench(L=24, width=12, depth=12),
    "periodic": lambda: cases.wavy_periodic(N=24),
    "sphere2d": lambda: cases.sphere2d(),
}

ADVECT = [
    ("sphere", 0, 0, cases.VEL_EGG), ("sphere", 1, 1, cases.VEL_EGG), ("sphere", 2, 0, cases.VEL_EGG),
    ("sphere", 3, 2, cases.VEL_EGG), ("sphere", 5, 1, cases.VEL_EGG), ("sphere", 6, 0, cases.VEL_EGG),
    ("sphere", 7, 0, cases.VEL_EGG), ("sphere", 8, 2, cases.VEL_EGG), ("sphere", 10, 2, cases.VEL_TRANSLATE_X),
    ("sphere", 11, 2, cases.VEL_TRANSLATE_X), ("trench", 11, 2, cases.VEL_ETCH_DIRECTIONAL),
    ("trench", 0, 0, cases.VEL_ETCH_DIRECTIONAL), ("periodic", 8, 2, cases.VEL_NORMAL_SQ),
    ("periodic", 11, 1, cases.VEL_NORMAL_SQ), ("sphere2d", 11, 2, dict(kind=0, b=1.0, vec=(0.5, 0.0, 0.0))),
    ("sphere2d", 8, 1, cases.VEL_GROW),
    # stencil local Lax-Friedrichs (enum 9)
    ("sphere", 9, 0, cases.VEL_EGG), ("sphere", 9, 2, cases.VEL_TRANSLATE_X),
    ("trench", 9, 0, cases.VEL_ETCH_DIRECTIONAL), ("periodic", 9, 1, cases.VEL_NORMAL_SQ),
    ("sphere2d", 9, 2, cases.VEL_GROW),
]


@pytest.mark.parametrize("geom,spatial,temporal,vel", ADVECT)
def test_advect(geom, spatial, temporal, vel):
    o = GEOMS[geom]()
    e = to_emul(o)
    p = po.make_params(spatial=spatial, temporal=temporal, cfl=0.4, time=0.8 if geom != "trench" else 2.0, vel=vel)
    so, ro = po.advect(o, p)
    se, re = pe.advect(e, p)
    assert so == 0 and se == 0
    assert ro["steps"] == re["steps"] and ro["advected_time"] == re["advected_time"]
    assert ro["point_updates"] == re["point_updates"]
    assert_same(o, e, f"advect {geom} {spatial} {temporal}")


def test_frame_fastdiv_exact():
    """Division by a run-time constant used to decode frame positions:
    exact against integer division on edge and random dividends."""
    rng = np.random.default_rng(9)
    divisors = [1, 2, 3, 5, 7, 16, 17, 255, 257, 641, 1025, 4097, 65537, 1 << 20, (1 << 20) + 3]
    divisors += list(rng.integers(1, 1 << 22, 20))
    for d in divisors:
        d = int(d)
        k = np.arange(0, (1 << 32) // d, max(1, ((1 << 32) // d) // 5000), dtype=np.uint64)
        n = np.concatenate([k * d, k * d + d - 1, k * d + 1, rng.integers(0, 1 << 32, 20000, dtype=np.uint64),
                            np.array([0, 1, (1 << 32) - 1, (1 << 32) - 2], np.uint64)])
        n = n[n < (1 << 32)].astype(np.uint32)
        assert np.array_equal(pe.fastdiv(d, n), n // np.uint32(d)), d


def _naive_deposit(x, m):
    out, j = 0, 0
    for b in range(64):
        if (m >> b) & 1:
            if (x >> j) & 1:
                out |= 1 << b
            j += 1
    return out


def test_bit_deposit_and_select():
    """Bit deposit / select of the index kernels against bit-by-bit loops."""
    rng = np.random.default_rng(10)
    full = (1 << 64) - 1
    ms = [0, full, 1, 1 << 63, 0x5555555555555555, 0xF0F0F0F0F0F0F0F0] + \
        [int(v) for v in rng.integers(0, 1 << 63, 300, dtype=np.uint64)] + \
        [int(v) & int(u) for v, u in zip(rng.integers(0, 1 << 63, 100, dtype=np.uint64),
                                         rng.integers(0, 1 << 63, 100, dtype=np.uint64))]
    xs = [int(v) | (int(u) << 32) for v, u in zip(rng.integers(0, 1 << 32, len(ms), dtype=np.uint64),
                                                   rng.integers(0, 1 << 32, len(ms), dtype=np.uint64))]
    xs[1] = full
    got = pe.deposit(np.array(xs, np.uint64), np.array(ms, np.uint64))
    for x, m, g in zip(xs, ms, got):
        assert int(g) == _naive_deposit(x, m), (hex(x), hex(m))
    ws, ks, want = [], [], []
    for m in ms:
        bits = [b for b in range(64) if (m >> b) & 1]
        for k in range(0, len(bits), 7):
            ws.append(m)
            ks.append(k)
            want.append(bits[k])
    assert np.array_equal(pe.select_bit(np.array(ws, np.uint64), np.array(ks, np.int32)), np.array(want))
