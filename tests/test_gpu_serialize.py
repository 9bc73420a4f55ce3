"""Domain::serialize / deserialize and the .lvst Writer / Reader
(lsDomain.hpp:188-268, lsWriter.hpp, lsReader.hpp) through the C ABI.

The envelope follows the reference byte for byte: "lsDomain", the version
byte 0, ..., the width as uint32 and the point-data flag byte.  The grid,
run-list and point-data sections in between are ViennaHRLE's and
ViennaCore's in the reference, which the reference tree does not contain:
their layout here is the library's own, so byte compatibility with
reference-written .lvst files is parity unpinned; these tests pin the
round trip (bit-identical band, width, grid and point data) and the
envelope."""
import struct
import warnings

import numpy as np
import pytest

import cases
import pyoracle as po
from test_gpu_parity import assert_same, say, to_gpu

pytestmark = pytest.mark.gpu


def _runs_equal(a, b):
    ia, va = a.get_runs()
    ib, vb = b.get_runs()
    return np.array_equal(ia, ib) and np.array_equal(va.view(np.int64), vb.view(np.int64))


def _with_point_data(g):
    n = g.getNumberOfPoints()
    pd = g.getPointData()
    rng = np.random.default_rng(3)
    pd.insertNextScalarData(rng.normal(size=n), "Material")
    pd.insertNextVectorData(rng.normal(size=(n, 3)), "Velocities")
    return g


def _fresh(D):
    import viennals_amd as vls
    return vls.Domain.from_index_bounds(D, [-4] * D, [4] * D, [cases.REF] * D, 0.5)


def test_round_trip_trench_with_point_data():
    say("serialize: trench with point data, round trip")
    g = _with_point_data(to_gpu(cases.trench(L=24, width=10, depth=8)))
    data = g.serialize()
    h = _fresh(3)
    h.deserialize(data)
    assert _runs_equal(g, h)
    assert h.getLevelSetWidth() == g.getLevelSetWidth()
    assert (h.gmin, h.gmax, h.bc, h.gridDelta) == (g.gmin, g.gmax, g.bc, g.gridDelta)
    for get in ("getScalarData", "getVectorData"):
        label = "Material" if get == "getScalarData" else "Velocities"
        a = np.asarray(getattr(g.getPointData(), get)(label))
        b = np.asarray(getattr(h.getPointData(), get)(label))
        assert np.array_equal(a.view(np.int64), b.view(np.int64))
    # serialising the copy gives the same bytes
    assert h.serialize() == data


def test_envelope_bytes():
    """"lsDomain", version 0, ..., width (uint32) and the point-data flag at
    the end (lsDomain.hpp:188-215)."""
    say("serialize: envelope bytes")
    o = cases.sphere2d(r=5.0)
    g = to_gpu(o)
    data = g.serialize()
    assert data[:8] == b"lsDomain" and data[8] == 0
    idx, val = g.get_runs()
    D, n = 2, len(val)
    grid = 8 + 1 + D * 9 + 8
    runs = 8 + 8 + n * (4 * D + 8)
    assert len(data) == 9 + grid + runs + 4 + 1
    assert struct.unpack_from("<I", data, 9 + grid + runs)[0] == o.width
    assert data[-1] == 0  # no point data
    assert struct.unpack_from("<Q", data, 9 + grid + 8)[0] == n


def test_writer_reader_file_and_advect(tmp_path):
    import viennals_amd as vls
    say("serialize: Writer / Reader through a .lvst file, then Advect")
    o = cases.sphere(r=5.0, dx=0.5)
    g = to_gpu(o)
    name = tmp_path / "sphere"
    w = vls.Writer(g, str(name))
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter("always")
        w.apply()
    assert any("appending" in str(r.message) for r in rec)
    assert (tmp_path / "sphere.lvst").exists()
    h = _fresh(3)
    vls.Reader(h, str(tmp_path / "sphere.lvst")).apply()
    assert _runs_equal(g, h)
    # the level set read back advects exactly like the oracle's
    p = po.make_params(spatial=11, temporal=2, cfl=0.4, time=0.5, vel=cases.VEL_GROW)
    assert po.advect(o, p)[0] == 0
    a = vls.Advect(h, vls.VelocityField(kind=0, b=1.0))
    a.setSpatialScheme(11)
    a.setTemporalScheme(2)
    a.setTimeStepRatio(0.4)
    a.setAdvectionTime(0.5)
    a.apply()
    assert_same(o, h, "advect after Reader")


def test_bad_streams_leave_the_level_set_unchanged():
    import viennals_amd as vls
    say("serialize: malformed streams")
    g = to_gpu(cases.sphere(r=4.0, dx=0.5))
    data = g.serialize()
    h = to_gpu(cases.sphere(r=3.0, dx=0.5))
    before = h.serialize()
    bad = [b"lsDomaiX" + data[8:],               # header
           data[:8] + b"\x01" + data[9:],        # newer version
           data[: len(data) // 2],               # cut short
           data[:-1]]                            # flag byte missing
    for b in bad:
        with pytest.raises(vls.LsgError):
            h.deserialize(b)
        assert h.serialize() == before
    with pytest.raises(vls.LsgError, match="2D level set"):
        h.deserialize(to_gpu(cases.sphere2d(r=4.0)).serialize())
    assert h.serialize() == before


def test_stream_without_point_data_keeps_point_data():
    """lsDomain.hpp:262-266 clears and reads point data only when the stream
    has some: a level set keeps its arrays when it reads one without."""
    say("serialize: point data kept when the stream has none")
    g = to_gpu(cases.sphere(r=4.0, dx=0.5))
    h = _with_point_data(to_gpu(cases.sphere(r=4.0, dx=0.5)))
    h.deserialize(g.serialize())
    assert _runs_equal(g, h)
    assert h.getPointData().getScalarDataSize() == 1
    assert h.getPointData().getVectorDataSize() == 1
